// context.hip -- tr_ctx: the device context behind the C ABI, replacing
// CudaContext (CudaContext.h:9-50, CudaContext.cpp:24-198).
//
// Differences from the reference that a caller can observe are none by
// design: vectors keep the reference's [vertex][aslm] layout at the copy
// boundary, norm/dot exclude the padding lanes, deviceAlloc zero-fills, and
// every call is ordered on one stream.  Internally vectors live renumbered
// and unpadded ([vertex][slm]) and the copy calls convert.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <set>
#include <vector>

#include "ctx.h"
#include "internal.h"
#include "kernels.h"
#include "cheb.h"
#include "solver.h"
#include "stencil.h"

namespace tr {

// The matrix-core schedule (kernels_mfma.hip): 16 vertices per wavefront.
LaneMap mfma_map(uint32_t slm, int wpb, int mode) {
    LaneMap m;
    m.variant = 4;
    m.WPB = wpb;
    m.J = mfma_J(slm, mode == 0 ? 8 : 4);
    m.L = 4;
    m.VPW = 16;
    m.mfma_mode = mode;
    return m;
}

// The lane mappings worth trying: J must divide slm with slm/J <= 64 lanes.
std::vector<LaneMap> map_candidates(uint32_t slm, int maxJ) {
    static const int Js[] = {1, 2, 3, 4, 5, 6, 7, 11};
    std::vector<LaneMap> out;
    bool any = false;
    for (int J : Js)
        if (J <= maxJ && slm % J == 0 && slm / J <= 64) any = true;
    if (!any) maxJ = 11;
    for (int J : Js) {
        if (J > maxJ || slm % J) continue;
        int L = (int)(slm / J);
        if (L > 64) continue;
        LaneMap m;
        m.J = J;
        m.L = L;
        m.VPW = 64 / L;
        if ((double)m.VPW * L / 64.0 < 0.5) continue;
        out.push_back(m);
    }
    return out;
}

// Picks the lane mapping: J must divide slm with slm/J <= 64 lanes; prefer
// the best lane utilisation, then the smaller J (fewer registers).
LaneMap choose_map(uint32_t slm, int maxJ) {
    static const int Js[] = {1, 2, 3, 4, 5, 6, 7, 11};
    LaneMap best;
    double bu = -1;
    // past maxJ (registers) only when no smaller J splits slm into <= 64
    // lanes (91 harmonics: J = 7)
    bool any = false;
    for (int J : Js)
        if (J <= maxJ && slm % J == 0 && slm / J <= 64) any = true;
    if (!any) maxJ = 11;
    for (int J : Js) {
        if (J > maxJ || slm % J) continue;
        int L = (int)(slm / J);
        if (L > 64) continue;
        int VPW = 64 / L;
        double u = (double)VPW * L / 64.0;
        if (u > bu + 1e-9) {
            bu = u;
            best.J = J;
            best.L = L;
            best.VPW = VPW;
        }
    }
    return best;
}

// Whether the matrix-core schedule covers this operator: ceil(slm/4) in
// {2, 4, 7}, exactly 15 or 16 volume slots, at most TR_WBMAX boundary slots.
static bool v4_ok(const HostOperator &op) {
    // 32-bit buffer offsets: the vector, the neighbour table and one slot's
    // coefficient records (FP64 sizes; a descriptor per slot) each stay
    // below 4 GiB
    const double lim = 4294967295.0;
    if ((double)op.nP * op.slm * 8 >= lim || (double)op.W * op.nP * 4 >= lim ||
        (double)op.nP * rec_stride<double>((int)op.NC) * 8 >= lim || (double)op.WB * op.nB * 4 * 8 >= lim)
        return false;
    return mfma_supported(op.slm, 0) && (op.W == 15 || op.W == 16) && op.WB <= TR_WBMAX &&
           (op.NC == 7 || op.NC == 10);
}

// Whether the buffer-load VALU schedule (variant 3) covers this operator
// with elem-byte values: at most 16 volume slots, and the vector, the
// neighbour table and one slot's coefficient records each below 4 GiB
// (32-bit buffer offsets, one descriptor per slot).
static bool v3_ok(const HostOperator &op, size_t elem) {
    const double lim = 4294967295.0;
    const int rs = elem == 8 ? rec_stride<double>((int)op.NC) : rec_stride<float>((int)op.NC);
    return op.W <= 16 && (double)op.nP * op.slm * elem < lim && (double)op.W * op.nP * 4 < lim &&
           (double)op.nP * rs * elem < lim;
}

// the VALU variant of a default schedule: 1 (plain loads) for wide
// stencils, 2 where the buffer-load schedule's offsets would not fit
static int valu_variant(const HostOperator &op, size_t elem, int v) {
    if (op.W > 16) return 1;
    return v == 3 && !v3_ok(op, elem) ? 2 : v;
}

// Whether the K-blocked matrix-core schedule (variant 5, FP32 operator)
// covers this operator: 33..80 harmonics, 15 or 16 volume slots.
static bool v5_ok(const HostOperator &op) {
    const double lim = 4294967295.0;
    if ((double)op.nP * op.slm * 4 >= lim || (double)op.W * op.nP * 4 >= lim ||
        (double)op.W * op.nP * rec_stride<float>((int)op.NC) * 4 >= lim || (double)op.WB * op.nB * 4 * 4 >= lim)
        return false;
    return mfma5_supported(op.slm) && (op.W == 15 || op.W == 16) && op.WB <= TR_WBMAX &&
           (op.NC == 7 || op.NC == 10);
}

static uint16_t to_bf16(float x);
static float from_bf16(uint16_t h);

// B blocks of variant 5: [section][harmonic block][channel][tile][lane], 4
// BF16 per lane (K index 4q + u of the block, output column 16t + l % 16)
// Harmonics of the internal (lm) order grouped into the four reflection-
// parity classes of the real even harmonics, cos(m phi) with m even, cos
// with m odd, sin(|m| phi) with |m| even, sin with |m| odd (classes by the
// mirror parities in x and y; z follows for even l), each in lm order.
// Omega_a Omega_b has definite mirror parities, so every coupling matrix
// maps a class onto one class: the diagonal channels (mass, xx, yy, zz,
// and the boundary's |Omega_n|) each onto itself, xy onto (0 <-> 2, 1 <->
// 3), xz onto (0 <-> 1, 2 <-> 3), yz onto (0 <-> 3, 1 <-> 2).
static std::vector<uint16_t> parity_class_order(uint32_t slm, std::vector<uint32_t> &sizes) {
    std::vector<uint16_t> order;
    sizes.assign(4, 0);
    for (int cls = 0; cls < 4; cls++)
        for (int l = 0, lm = 0; (uint32_t)lm < slm; l += 2)
            for (int m = -l; m <= l; m++, lm++) {
                const int am = m < 0 ? -m : m;
                const int k = (m >= 0 ? 0 : 2) + (am & 1);
                if (k == cls) {
                    order.push_back((uint16_t)lm);
                    sizes[cls]++;
                }
            }
    return order;
}

// B blocks of variant 5 and their metadata for the FP32 vectors' order hp
// (hp[k] = internal harmonic at position k) split into the given blocks:
// per input block kb, channel c and output block t the 16 x 16 block
// (lane l holds K rows 4(l/16) .. +3 of column l % 16, BF16), stored only
// when some coupling in it is nonzero.  Layout: volume hi blocks (kb, c, t
// order), then volume lo blocks when lo, then the boundary's likewise.
static std::vector<unsigned char> mfma_tables5(const HostOperator &op, const std::vector<uint16_t> &hp,
                                               const std::vector<uint32_t> &h0, const std::vector<uint32_t> &w,
                                               bool lo, Blk5 &b5) {
    const uint32_t slm = op.slm, NT = (uint32_t)h0.size();
    auto mat = [&](int sec, int c, uint32_t n, uint32_t m) -> double {
        if (sec == 1) return op.onc[((size_t)c * slm + n) * slm + m];
        if (c == 0) return n == m ? op.massfac[n] : 0.0;
        return op.omc[((size_t)(c - 1) * slm + n) * slm + m];
    };
    b5 = Blk5();
    b5.nb = NT;
    b5.nsplitB = lo ? 2 : 1;
    for (uint32_t k = 0; k < NT; k++) {
        b5.h0[k] = h0[k];
        b5.w[k] = w[k];
    }
    std::vector<unsigned char> out;
    for (int sec = 0; sec < 2; sec++) {
        const int nch = sec == 0 ? (int)op.NC : 3;
        // couplings below 1e-12 of the section's largest are the rounding
        // residues of the quadrature (the exact values vanish by parity):
        // zero, so that their blocks are not stored
        double big = 0;
        for (int c = 0; c < nch; c++)
            for (uint32_t n = 0; n < slm; n++)
                for (uint32_t m = 0; m < slm; m++) big = std::max(big, std::fabs(mat(sec, c, n, m)));
        const double tiny = 1e-12 * big;
        std::vector<unsigned char> hi_part, lo_part;
        uint32_t count = 0;
        for (uint32_t kb = 0; kb < NT; kb++) {
            uint64_t mask = 0;
            const uint32_t first = count;
            for (int c = 0; c < nch; c++)
                for (uint32_t tt = 0; tt < NT; tt++) {
                    unsigned char bh[64 * 8], bl[64 * 8];
                    bool nz = false;
                    for (int l = 0; l < 64; l++) {
                        const uint32_t col = (uint32_t)(l % 16), q = (uint32_t)(l / 16);
                        uint16_t h[4], g[4];
                        for (int u = 0; u < 4; u++) {
                            const uint32_t kk = 4u * q + (uint32_t)u;
                            double x = 0.0;
                            if (col < w[tt] && kk < w[kb]) x = mat(sec, c, hp[h0[tt] + col], hp[h0[kb] + kk]);
                            if (std::fabs(x) <= tiny) x = 0.0;
                            nz = nz || x != 0.0;
                            h[u] = to_bf16((float)x);
                            g[u] = to_bf16((float)x - from_bf16(h[u]));
                        }
                        std::memcpy(bh + l * 8, h, 8);
                        std::memcpy(bl + l * 8, g, 8);
                    }
                    if (!nz) continue;
                    mask |= 1ull << (c * NT + tt);
                    hi_part.insert(hi_part.end(), bh, bh + sizeof bh);
                    lo_part.insert(lo_part.end(), bl, bl + sizeof bl);
                    count++;
                }
            if (sec == 0) {
                b5.vmask[kb] = mask;
                b5.voff[kb] = first;
            } else {
                b5.bmask[kb] = mask;
                b5.boff[kb] = first;
            }
        }
        (sec == 0 ? b5.nvol : b5.nbnd) = count;
        out.insert(out.end(), hi_part.begin(), hi_part.end());
        if (lo) out.insert(out.end(), lo_part.begin(), lo_part.end());
    }
    return out;
}

// The B blocks of the static class layout sl (kernels.h) for the class
// order hp: every pair of the layout in cl_vol_index / cl_bnd_index order,
// zero blocks included, hi parts then lo parts per section (volume, then
// boundary), or (f32) the FP32 blocks per section.  False when a coupling
// outside the layout's pairs is not zero, i.e. the layout does not
// describe this operator.
static bool mfma_tables5s(const HostOperator &op, const std::vector<uint16_t> &hp, int sl, Blk5 &b5,
                          std::vector<unsigned char> &out, bool f32) {
    const uint32_t slm = op.slm;
    const int NT = cl_nblk(sl), NC = (int)op.NC;
    if (slm != cl_slm(sl) || hp.size() != slm) return false;
    auto mat = [&](int sec, int c, uint32_t n, uint32_t m) -> double {
        if (sec == 1) return op.onc[((size_t)c * slm + n) * slm + m];
        if (c == 0) return n == m ? op.massfac[n] : 0.0;
        return op.omc[((size_t)(c - 1) * slm + n) * slm + m];
    };
    b5 = Blk5();
    b5.nb = (uint32_t)NT;
    b5.nsplitB = 2;
    b5.stat = (uint32_t)sl;
    b5.f32 = f32 ? 1u : 0u;
    for (int k = 0; k < NT; k++) {
        b5.h0[k] = (uint32_t)cl_h0(sl, k);
        b5.w[k] = (uint32_t)cl_w(sl, k);
    }
    out.clear();
    for (int sec = 0; sec < 2; sec++) {
        const int nch = sec == 0 ? NC : 3;
        double big = 0;
        for (int c = 0; c < nch; c++)
            for (uint32_t n = 0; n < slm; n++)
                for (uint32_t m = 0; m < slm; m++) big = std::max(big, std::fabs(mat(sec, c, n, m)));
        const double tiny = 1e-12 * big;
        std::vector<unsigned char> hi_part, lo_part;
        uint32_t npairs = 0;
        for (int kb = 0; kb < NT; kb++)
            for (int c = 0; c < nch; c++)
                for (int t = 0; t < NT; t++) {
                    unsigned char bh[64 * 8], bl[64 * 8], bf[64 * 16];
                    bool nz = false;
                    for (int l = 0; l < 64; l++) {
                        const uint32_t col = (uint32_t)(l % 16), q = (uint32_t)(l / 16);
                        uint16_t h[4], g[4];
                        float xf[4];
                        for (int u = 0; u < 4; u++) {
                            const uint32_t kk = 4u * q + (uint32_t)u;
                            double x = 0.0;
                            if (col < b5.w[t] && kk < b5.w[kb]) x = mat(sec, c, hp[b5.h0[t] + col], hp[b5.h0[kb] + kk]);
                            if (std::fabs(x) <= tiny) x = 0.0;
                            nz = nz || x != 0.0;
                            h[u] = to_bf16((float)x);
                            g[u] = to_bf16((float)x - from_bf16(h[u]));
                            xf[u] = (float)x;
                        }
                        std::memcpy(bh + l * 8, h, 8);
                        std::memcpy(bl + l * 8, g, 8);
                        // lane l, element u: B[4 (l / 16) + u][l % 16], slice u
                        // of the four FP32 products
                        std::memcpy(bf + l * 16, xf, 16);
                    }
                    const bool pair = sec == 0 ? cl_vol_pair(sl, kb, c, t) : cl_bnd_pair(sl, kb, t);
                    if (!pair) {
                        if (nz) return false;
                        continue;
                    }
                    npairs++;
                    if (f32) {
                        hi_part.insert(hi_part.end(), bf, bf + sizeof bf);
                    } else {
                        hi_part.insert(hi_part.end(), bh, bh + sizeof bh);
                        lo_part.insert(lo_part.end(), bl, bl + sizeof bl);
                    }
                }
        (sec == 0 ? b5.nvol : b5.nbnd) = npairs;
        out.insert(out.end(), hi_part.begin(), hi_part.end());
        out.insert(out.end(), lo_part.begin(), lo_part.end());
    }
    return b5.nvol == (uint32_t)cl_vol_count(sl, NC) && b5.nbnd == (uint32_t)cl_bnd_count(sl);
}

// The FP64 B blocks of variant 5 for the class order hp split into blocks
// (h0, w): per input block kb, channel c and output block t whose couplings
// are not all zero, 4 K steps x 64 lanes of doubles (lane l of step st: the
// coupling of output hp[h0_t + l % 16] with input hp[h0_kb + mfma_harm(2,
// l / 16, st)], the operand layout of v_mfma_f64_16x16x4f64 after the move
// from the gather layout).  Volume blocks, then the boundary's.
static std::vector<unsigned char> mfma_tables5d(const HostOperator &op, const std::vector<uint16_t> &hp,
                                                const std::vector<uint32_t> &h0, const std::vector<uint32_t> &w,
                                                Blk5 &b5) {
    const uint32_t slm = op.slm, NT = (uint32_t)h0.size();
    auto mat = [&](int sec, int c, uint32_t n, uint32_t m) -> double {
        if (sec == 1) return op.onc[((size_t)c * slm + n) * slm + m];
        if (c == 0) return n == m ? op.massfac[n] : 0.0;
        return op.omc[((size_t)(c - 1) * slm + n) * slm + m];
    };
    b5 = Blk5();
    b5.nb = NT;
    b5.nsplitB = 1;
    for (uint32_t k = 0; k < NT; k++) {
        b5.h0[k] = h0[k];
        b5.w[k] = w[k];
    }
    std::vector<unsigned char> out;
    for (int sec = 0; sec < 2; sec++) {
        const int nch = sec == 0 ? (int)op.NC : 3;
        // the same threshold as the FP32 blocks: couplings below 1e-12 of
        // the largest are rounding residues of entries that vanish by parity
        double big = 0;
        for (int c = 0; c < nch; c++)
            for (uint32_t n = 0; n < slm; n++)
                for (uint32_t m = 0; m < slm; m++) big = std::max(big, std::fabs(mat(sec, c, n, m)));
        const double tiny = 1e-12 * big;
        uint32_t count = 0;
        for (uint32_t kb = 0; kb < NT; kb++) {
            uint64_t mask = 0;
            const uint32_t first = count;
            for (int c = 0; c < nch; c++)
                for (uint32_t tt = 0; tt < NT; tt++) {
                    double blk[4 * 64];
                    bool nz = false;
                    for (int st = 0; st < 4; st++)
                        for (int l = 0; l < 64; l++) {
                            const uint32_t col = (uint32_t)(l % 16), kk = mfma_harm(2, l / 16, st);
                            double x = 0.0;
                            if (col < w[tt] && kk < w[kb]) x = mat(sec, c, hp[h0[tt] + col], hp[h0[kb] + kk]);
                            if (std::fabs(x) <= tiny) x = 0.0;
                            nz = nz || x != 0.0;
                            blk[st * 64 + l] = x;
                        }
                    if (!nz) continue;
                    mask |= 1ull << (c * NT + tt);
                    const unsigned char *bp = reinterpret_cast<const unsigned char *>(blk);
                    out.insert(out.end(), bp, bp + sizeof blk);
                    count++;
                }
            if (sec == 0) {
                b5.vmask[kb] = mask;
                b5.voff[kb] = first;
            } else {
                b5.bmask[kb] = mask;
                b5.boff[kb] = first;
            }
        }
        (sec == 0 ? b5.nvol : b5.nbnd) = count;
    }
    return out;
}

// The parity-class blocks of an operator with 33..80 harmonics (at most 16
// harmonics per block inside each class), or false when the classes do not
// apply: batches, reflecting planes and faces off the axes index harmonics
// in the internal order, and the blocks must number 3 to 5.
static bool class_blocks(const tr_ctx *c, std::vector<uint16_t> &hp, std::vector<uint32_t> &h0,
                         std::vector<uint32_t> &w) {
    const HostOperator &op = c->op;
    std::vector<uint32_t> sizes;
    hp = parity_class_order(op.slm, sizes);
    h0.clear();
    w.clear();
    for (int cls = 0, base = 0; cls < 4; base += (int)sizes[cls], cls++)
        for (uint32_t k = 0; k < sizes[cls]; k += 16) {
            h0.push_back((uint32_t)base + k);
            w.push_back(std::min<uint32_t>(16, sizes[cls] - k));
        }
    return mfma5_supported(op.slm) && c->voff.empty() && op.gvert.empty() && op.mask.empty() && h0.size() >= 3 &&
           h0.size() <= 5;
}

// FP64 variant 5 for this context where it applies: the class blocks, 15 or
// 16 volume slots, FP64 arrays within 32-bit buffer offsets, and the stored
// volume blocks within the LDS of one workgroup.  Leaves d_tab5d null
// otherwise.
static int build_tables5d(tr_ctx *c, int device) {
    const HostOperator &op = c->op;
    std::vector<uint16_t> hp;
    std::vector<uint32_t> h0, w;
    if (!class_blocks(c, hp, h0, w)) return TR_OK;
    const double lim = 4294967295.0;
    if ((double)op.nP * op.slm * 8 >= lim || (double)op.W * op.nP * 4 >= lim ||
        (double)op.W * op.nP * rec_stride<double>((int)op.NC) * 8 >= lim || (double)op.WB * op.nB * 4 * 8 >= lim ||
        (op.W != 15 && op.W != 16) || op.WB > TR_WBMAX || (op.NC != 7 && op.NC != 10))
        return TR_OK;
    Blk5 b5;
    std::vector<unsigned char> tab = mfma_tables5d(op, hp, h0, w, b5);
    int maxlds = 65536;
    hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
    if (apply5d_lds_bytes(b5) + 64 > (size_t)maxlds) return TR_OK;
    int rc;
    if ((rc = upload(c, (unsigned char **)&c->d_tab5d, tab)) || (rc = upload(c, &c->d_hp64, hp))) return rc;
    c->b5d = b5;
    c->hp64 = hp;
    return TR_OK;
}

// Variant 5's FP32 order and blocks for this context: the parity classes
// (blocks of at most 16 inside each class) for a single problem without
// reflecting planes or faces off the axes (their kernels index harmonics in
// the internal order), when that gives 3 to 5 blocks; else the internal
// order in blocks of 16.  The lo parts of B are kept when the stored
// blocks take at most 80 KB with them (two workgroups per CU).
static int build_tables5(tr_ctx *c) {
    const HostOperator &op = c->op;
    const uint32_t slm = op.slm;
    std::vector<uint32_t> h0, w;
    std::vector<uint16_t> hp;
    const bool classes = class_blocks(c, hp, h0, w);
    if (!classes) {
        hp.resize(slm);
        for (uint32_t k = 0; k < slm; k++) hp[k] = (uint16_t)k;
        h0.clear();
        w.clear();
        for (uint32_t k = 0; k < slm; k += 16) {
            h0.push_back(k);
            w.push_back(std::min<uint32_t>(16, slm - k));
        }
    }
    Blk5 b5;
    std::vector<unsigned char> tab;
    // the static class layout (straight-line contraction) where it applies
    const int sl = op.slm == cl_slm(5) ? 5 : (op.slm == cl_slm(4) ? 4 : 0);
    bool stat = classes && c->tab5_static && sl && (op.NC == 7 || op.NC == 10) && op.W == 15 &&
                h0.size() == (size_t)cl_nblk(sl);
    for (size_t k = 0; stat && k < h0.size(); k++)
        stat = h0[k] == (uint32_t)cl_h0(sl, (int)k) && w[k] == (uint32_t)cl_w(sl, (int)k);
    if (!(stat && mfma_tables5s(op, hp, sl, b5, tab, c->ang_prec == 4))) {
        tab = mfma_tables5(op, hp, h0, w, true, b5);
        if ((size_t)b5.nvol * 1024 > 80 * 1024) tab = mfma_tables5(op, hp, h0, w, false, b5);
    }
    if (c->f_tab5) {
        // a rebuild (the layout switched): the old table goes once no
        // queued launch can read it
        HIPCHECK(hipStreamSynchronize(c->stream));
        c->owned.erase(std::remove(c->owned.begin(), c->owned.end(), c->f_tab5), c->owned.end());
        HIPCHECK(hipFree(c->f_tab5));
        c->f_tab5 = nullptr;
    }
    int rc = upload(c, (unsigned char **)&c->f_tab5, tab);
    if (rc) return rc;
    c->b5 = b5;
    if (classes) {
        c->hp32 = hp;
        if ((rc = upload(c, &c->d_hp32, hp))) return rc;
    } else {
        c->hp32.clear();
        c->d_hp32 = nullptr;
    }
    return TR_OK;
}

// The K-blocked matrix-core schedule with its persistent grids.
static LaneMap mfma5_map_ctx(const tr_ctx *c, int wpb) {
    LaneMap m;
    m.variant = 5;
    m.WPB = wpb;
    m.J = 4;
    m.L = 4;
    m.VPW = 16;
    m.mfma_mode = 2;
    const int NC = (int)c->op.NC;
    m.grid = apply5_grid(c->args32(m), NC, wpb, 0);
    m.grid_dot = apply5_grid(c->args32(m), NC, wpb, 1);
    m.grid_cheb = apply5_grid(c->args32(m), NC, wpb, 2);
    return m;
}

// The FP64 K-blocked schedule (variant 5) with its persistent grids; the
// scratch vector of internal-order inputs is allocated on first use.
static int mfma5d_map_ctx(tr_ctx *c, int wpb, LaneMap &m) {
    if (!c->d_tab5d) return fail(TR_ERR_ARG, "FP64 K-blocked matrix-core schedule unsupported here");
    if (!c->d_cls) {
        HIPCHECK(hipMalloc((void **)&c->d_cls, sizeof(double) * std::max<size_t>(c->nint(), 1)));
        c->owned.push_back(c->d_cls);
    }
    m = LaneMap();
    m.variant = 5;
    m.WPB = wpb;
    m.J = 4;
    m.L = 4;
    m.VPW = 16;
    m.mfma_mode = 0;
    const int NC = (int)c->op.NC;
    m.grid = apply5d_grid(c->args(m), NC, wpb, 0);
    m.grid_dot = apply5d_grid(c->args(m), NC, wpb, 1);
    m.grid_cheb = apply5d_grid(c->args(m), NC, wpb, 2);
    return TR_OK;
}

// The matrix-core schedule of this context with its persistent grid sized
// from the kernel's occupancy.
static LaneMap mfma_map_ctx(const tr_ctx *c, int wpb, int mode) {
    LaneMap m = mfma_map(c->op.slm, wpb, mode);
    const int NC = (int)c->op.NC;
    for (int epi : {0, 1, 2}) {
        uint32_t g;
        if (mode == 0)
            g = apply4_grid<double, 0>(c->args(m), NC, wpb, epi);
        else if (mode == 1)
            g = apply4_grid<float, 1>(c->args32(m), NC, wpb, epi);
        else
            g = apply4_grid<float, 2>(c->args32(m), NC, wpb, epi);
        (epi == 0 ? m.grid : (epi == 1 ? m.grid_dot : m.grid_cheb)) = g;
    }
    return m;
}

static size_t lds_need(const LaneMap &m, const HostOperator &op, size_t elem, int nsplit = 3) {
    // variant 5 stores at most 250 B blocks (128 KB) hi only, or at most
    // 80 KB with lo parts (build_tables5): always within the LDS
    if (m.variant == 5) return 64;
    if (m.variant == 4) return apply4_lds_bytes(elem == 8 ? 0 : m.mfma_mode, (int)op.NC, op.slm, nsplit) + 64;
    if (m.variant == 3) return apply3_lds_bytes(m, (int)op.NC, op.slm, op.arow, op.brow, elem);
    return m.variant == 2 ? apply2_lds_bytes(m, (int)op.NC, op.slm, op.arow, op.brow, elem)
                          : apply_lds_bytes(m, (int)op.NC, op.slm, elem);
}

// Fits a VALU schedule into the device's LDS: fewer wavefronts per
// workgroup first, then the plain-load schedule (variant 1), which stages
// no angular tables (at 91 or 120 harmonics the tables of variants 2 and 3
// alone exceed the LDS).  False when nothing fits.
static bool fit_lds(LaneMap &m, const HostOperator &op, size_t elem, int device) {
    int maxlds = 65536;
    hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
    while (m.WPB > 1 && lds_need(m, op, elem) > (size_t)maxlds) m.WPB /= 2;
    if (lds_need(m, op, elem) > (size_t)maxlds && m.variant >= 2 && m.variant <= 3) {
        m.variant = 1;
        m.WPB = 4;
        while (m.WPB > 1 && lds_need(m, op, elem) > (size_t)maxlds) m.WPB /= 2;
    }
    return lds_need(m, op, elem) <= (size_t)maxlds;
}

// [W][nP][RS] coefficient records from the [W][NC][nP] layout
template <typename T>
static std::vector<T> coef_records(const HostOperator &op) {
    const int RS = rec_stride<T>((int)op.NC);
    std::vector<T> rec((size_t)op.W * op.nP * RS, (T)0);
    for (uint32_t s = 0; s < op.W; s++)
        for (uint32_t v = 0; v < op.nP; v++)
            for (uint32_t c = 0; c < op.NC; c++)
                rec[((size_t)s * op.nP + v) * RS + c] = (T)op.coef[((size_t)s * op.NC + c) * op.nP + v];
    return rec;
}

// [WBP][nB][4] boundary coefficient records (3 channels, one pad) from the
// [WB][3][nB] layout, WBP = max(WB, TR_WBMAX): the matrix-core schedule
// walks TR_WBMAX slots, the ones past WB with zero coefficients
template <typename T>
static std::vector<T> bcoef_records(const HostOperator &op) {
    const uint32_t WBP = std::max<uint32_t>(op.WB, TR_WBMAX);
    std::vector<T> rec((size_t)WBP * op.nB * 4, (T)0);
    for (uint32_t s = 0; s < op.WB; s++)
        for (uint32_t b = 0; b < op.nB; b++)
            for (int n = 0; n < 3; n++)
                rec[((size_t)s * op.nB + b) * 4 + n] = (T)op.bcoef[((size_t)s * 3 + n) * op.nB + b];
    return rec;
}

// Order in which the matrix-core schedule walks its wavefront tiles (16
// consecutive vertices each): tiles holding vacuum-boundary vertices do the
// boundary gather as well, about twice the work of the others, so they are
// spread evenly through the order (each XCD's contiguous share of the
// persistent grid's work gets its share of them); both kinds keep their
// (Morton) order, so neighbouring entries stay spatially close.
static std::vector<uint32_t> tile_order(const HostOperator &op) {
    const uint32_t nt = (op.nP + 15) / 16;
    std::vector<uint32_t> bt, it, out;
    for (uint32_t t = 0; t < nt; t++) {
        bool b = false;
        for (uint32_t v = 16 * t; v < std::min(op.nP, 16 * t + 16); v++) b = b || op.bof[v] >= 0;
        (b ? bt : it).push_back(t);
    }
    out.reserve(nt);
    size_t ib = 0, ii = 0;
    for (uint32_t k = 0; k < nt; k++) {
        // boundary tiles due after k + 1 entries: floor((k + 1) nb / nt)
        const size_t due = (size_t)(k + 1) * bt.size() / nt;
        if (ib < due || ii == it.size())
            out.push_back(bt[ib++]);
        else
            out.push_back(it[ii++]);
    }
    return out;
}

static uint16_t to_bf16(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static float from_bf16(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

// MFMA B slices of the variant-4 schedule (kernels_mfma.hip): for channel c,
// K-step st, output tile t and lane l the coupling Mat_c[n][m] with
// n = 16t + l%16 and m = mfma_kharm(mode, J, l/16, st) (FP64/FP32, one value) or
// m = mfma_harm(E, l/16, 4st + u), u = 0..3 (BF16, four values), E = 2 for
// FP64 and 4 for FP32 (kernels.h); volume channels
// (mass = diag(massfac), then the NC-1 stiffness channels) first, then the 3
// boundary channels.  mode 0 FP64, 1 FP32, 2 BF16 (nsplit 3 appends the
// low parts: hi table, lo table, per section).
// [WBP][nB] boundary neighbours, WBP as in bcoef_records: the extra slots
// point at the boundary vertex itself
static std::vector<uint32_t> bnbr_padded(const HostOperator &op) {
    const uint32_t WBP = std::max<uint32_t>(op.WB, TR_WBMAX);
    std::vector<uint32_t> out((size_t)WBP * op.nB);
    for (uint32_t s = 0; s < WBP; s++)
        for (uint32_t b = 0; b < op.nB; b++) out[(size_t)s * op.nB + b] = s < op.WB ? op.bnbr[(size_t)s * op.nB + b] : op.bvert[b];
    return out;
}

static std::vector<unsigned char> mfma_tables(const HostOperator &op, int mode, int nsplit) {
    const uint32_t slm = op.slm;
    const int E = mode == 0 ? 2 : 4, J = mfma_J(slm, mode == 0 ? 8 : 4);
    const int S = mfma_steps(mode, J), NT = (int)((slm + 15) / 16);
    const int BW = mode == 2 ? 8 : (mode == 0 ? 8 : 4);
    auto mat = [&](int sec, int c, uint32_t n, uint32_t m) -> double {
        if (n >= slm || m >= slm) return 0.0;
        if (sec == 1) return op.onc[((size_t)c * slm + n) * slm + m];
        if (c == 0) return n == m ? op.massfac[n] : 0.0;
        return op.omc[((size_t)(c - 1) * slm + n) * slm + m];
    };
    std::vector<unsigned char> out;
    for (int sec = 0; sec < 2; sec++) {
        const int nch = sec == 0 ? (int)op.NC : 3;
        const size_t sz = (size_t)nch * S * NT * 64 * BW;
        for (int part = 0; part < (mode == 2 && nsplit > 1 ? 2 : 1); part++) {
            std::vector<unsigned char> t(sz, 0);
            for (int c = 0; c < nch; c++)
                for (int st = 0; st < S; st++)
                    for (int tt = 0; tt < NT; tt++)
                        for (int l = 0; l < 64; l++) {
                            unsigned char *p = &t[((((size_t)c * S + st) * NT + tt) * 64 + l) * BW];
                            const uint32_t n = 16u * tt + (l % 16), q = l / 16;
                            if (mode == 0) {
                                double v = mat(sec, c, n, mfma_kharm(0, J, (int)q, st));
                                std::memcpy(p, &v, 8);
                            } else if (mode == 1) {
                                float v = (float)mat(sec, c, n, mfma_kharm(1, J, (int)q, st));
                                std::memcpy(p, &v, 4);
                            } else {
                                uint16_t h[4];
                                for (int u = 0; u < 4; u++) {
                                    double v = (4 * st + u < J) ? mat(sec, c, n, mfma_harm(E, (int)q, 4 * st + u)) : 0.0;
                                    uint16_t hi = to_bf16((float)v);
                                    h[u] = part == 0 ? hi : to_bf16((float)(v - (double)from_bf16(hi)));
                                }
                                std::memcpy(p, h, 8);
                            }
                        }
            out.insert(out.end(), t.begin(), t.end());
        }
    }
    return out;
}

// contraction table records {value, byte offset into the staged vertex}:
// volume entries idx = (c-1)*slm + lm' sit in staging channel c, boundary
// entries idx = n*slm + lm' in channel NC + n
template <typename T>
static std::vector<unsigned char> table_records(const HostOperator &op) {
    const int TB = tab_rec_bytes<T>();
    const uint32_t slm = op.slm;
    std::vector<unsigned char> out((size_t)(op.arow + op.brow) * slm * TB, 0);
    auto put = [&](size_t row, uint32_t lm, double val, uint32_t off) {
        unsigned char *p = &out[(row * slm + lm) * TB];
        T v = (T)val;
        std::memcpy(p, &v, sizeof(T));
        std::memcpy(p + sizeof(T), &off, 4);
    };
    for (uint32_t e = 0; e < op.arow; e++)
        for (uint32_t lm = 0; lm < slm; lm++) {
            size_t q = (size_t)e * slm + lm;
            put(e, lm, op.aval[q], (uint32_t)((op.aidx[q] + slm) * sizeof(T)));
        }
    for (uint32_t e = 0; e < op.brow; e++)
        for (uint32_t lm = 0; lm < slm; lm++) {
            size_t q = (size_t)e * slm + lm;
            put(op.arow + e, lm, op.bval[q], (uint32_t)((op.NC * slm + op.bidx[q]) * sizeof(T)));
        }
    return out;
}


static int lhs(tr_ctx *c, const double *f, double *r) {
    HIPCHECK(launch_apply<double>(c->args(), (int)c->op.NC, c->map, f, r, nullptr, nullptr, c->stream));
    return TR_OK;
}

static int lhs_diag(tr_ctx *c, const double *f, double *r) {
    // the angular-diagonal operator has VALU schedules only
    HIPCHECK(launch_apply_diag<double>(c->args(c->vmap), (int)c->op.NC, c->vmap, f, r, nullptr, nullptr, c->stream));
    return TR_OK;
}

static double dot(tr_ctx *c, const double *x, const double *y) {
    hipError_t e = launch_dot(x, y, c->nint(), c->d_part, c->d_red, c->stream);
    double h = NAN;
    if (e == hipSuccess) e = hipMemcpyAsync(&h, c->d_red, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        fail(TR_ERR_HIP, std::string("dot: ") + hipGetErrorString(e));
        return NAN;
    }
    return h;
}

static int axpby(tr_ctx *c, double *x, double wx, const double *y, double wy) {
    HIPCHECK(launch_axpby(x, wx, y, wy, c->nint(), c->stream));
    return TR_OK;
}

// r = b - Ap without reading r first (0 * r keeps a NaN that r held)
static int residual(tr_ctx *c, double *r, const double *b, const double *Ap) {
    HIPCHECK(hipMemcpyAsync(r, b, sizeof(double) * c->nint(), hipMemcpyDeviceToDevice, c->stream));
    return axpby(c, r, 1, Ap, -1);
}

}  // namespace tr

using namespace tr;

extern "C" {

int tr_ctx_create(int device, const tr_mesh_view *mesh, const tr_angular_view *ang, int flags,
                  tr_ctx **out) {
    if (!mesh || !ang || !out) return fail(TR_ERR_ARG, "tr_ctx_create: null argument");
    if (flags & ~TR_CTX_GENERAL_NORMALS) return fail(TR_ERR_ARG, "tr_ctx_create: unknown flags");
    *out = nullptr;
    tr_ctx *c = new tr_ctx();
    c->device = device;
    c->op.general_normals = (flags & TR_CTX_GENERAL_NORMALS) != 0;
    int rc = assemble_operator(*mesh, *ang, c->op);
    if (rc) {
        delete c;
        return rc;
    }
    if ((rc = ctx_init(c, device))) return rc;
    *out = c;
    return TR_OK;
}

}  // extern "C"

int tr::ctx_init(tr_ctx *c, int device) {
    int rc;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess && !c->stream) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(TR_ERR_HIP, std::string("tr_ctx_create: ") + hipGetErrorString(e));
    }
    if ((rc = noalloc_init(device))) {  // this GPU's address of the allocation-free dot's scratch
        tr_ctx_destroy(c);
        return rc;
    }
    c->map = choose_map(c->op.slm, 4);
    c->map.variant = valu_variant(c->op, 8, c->map.variant);
    if (!fit_lds(c->map, c->op, 8, device)) {
        tr_ctx_destroy(c);
        return fail(TR_ERR_ARG, "tr_ctx_create: angular order too large for LDS staging");
    }
    c->vmap = c->map;
    const HostOperator &op = c->op;
    std::vector<double> stage;
    if ((rc = upload(c, &c->d_nbr, op.nbr)) || (rc = upload(c, &c->d_coef, op.coef)) ||
        (rc = upload(c, &c->d_bof, op.bof)) || (rc = upload(c, &c->d_bnbr, bnbr_padded(op))) ||
        (rc = upload(c, &c->d_bcoef, op.bcoef)) || (rc = upload(c, &c->d_aidx, op.aidx)) ||
        (rc = upload(c, &c->d_aval, op.aval)) || (rc = upload(c, &c->d_bidx, op.bidx)) ||
        (rc = upload(c, &c->d_bval, op.bval)) || (rc = upload(c, &c->d_massfac, op.massfac)) ||
        (rc = upload(c, &c->d_ddiag, op.ddiag)) || (rc = upload(c, &c->d_bdiag, op.bdiag)) ||
        (rc = upload(c, &c->d_rhs, op.rhs)) || (rc = upload(c, &c->d_perm, op.perm)) ||
        (rc = upload(c, &c->d_iperm, op.iperm)) || (rc = upload(c, &c->d_crec, coef_records<double>(op))) ||
        (rc = upload(c, &c->d_bcrec, bcoef_records<double>(op))) || (rc = upload(c, &c->d_tiles, tile_order(op))) ||
        (!op.mask.empty() && (rc = upload(c, &c->d_mask, op.mask))) ||
        (rc = upload(c, &c->d_mom_val, op.mom_val)) || (rc = upload(c, &c->d_mom_pos, op.mom_pos)) ||
        (!op.mf0.empty() && (rc = upload(c, &c->d_mf0, op.mf0))) ||
        (!op.vsum.empty() && (rc = upload(c, &c->d_vsum, op.vsum))) ||
        (!op.xyz.empty() && (rc = upload(c, &c->d_xyz, op.xyz))) ||
        (!op.gvert.empty() &&
         ((rc = upload(c, &c->d_gvert, op.gvert)) || (rc = upload(c, &c->d_gstart, op.gstart)) ||
          (rc = upload(c, &c->d_gface, op.gface)) || (rc = upload(c, &c->d_gpos, op.gpos)) ||
          (rc = upload(c, &c->d_gfv, op.gfv)) || (rc = upload(c, &c->d_garea, op.garea)) ||
          (rc = upload(c, &c->d_gR, op.gR)) || (rc = upload(c, &c->d_gOd, op.gOd)) ||
          (rc = upload(c, &c->d_gOz, op.gOz)))) ||
        (rc = upload(c, (unsigned char **)&c->d_tab, table_records<double>(op))) ||
        (mfma_supported(op.slm, 0) && (rc = upload(c, (unsigned char **)&c->d_tab4, mfma_tables(op, 0, 1)))) ||
        (rc = build_tables5d(c, device))) {
        tr_ctx_destroy(c);
        return rc;
    }
    e = hipMalloc((void **)&c->d_part, sizeof(double) * (red_partials() + 8));
    if (e == hipSuccess) {
        c->owned.push_back(c->d_part);
        c->d_red = c->d_part + red_partials();
        e = hipMalloc((void **)&c->d_stage, sizeof(double) * c->next());
    }
    if (e == hipSuccess) {
        c->owned.push_back(c->d_stage);
        e = hipStreamSynchronize(c->stream);
    }
    if (e != hipSuccess) {
        tr_ctx_destroy(c);
        return fail(TR_ERR_HIP, std::string("tr_ctx_create: ") + hipGetErrorString(e));
    }
    return TR_OK;
}

extern "C" {

int tr_ctx_destroy(tr_ctx *c) {
    if (!c) return TR_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    destroy_multilevel(c->mg);
    for (void *p : c->owned) hipFree(p);
    for (void *p : c->vectors) hipFree(p);
    for (void *p : c->raw) hipFree(p);
    if (c->stream && c->own_stream) hipStreamDestroy(c->stream);
    delete c;
    return TR_OK;
}

uint32_t tr_ctx_N(const tr_ctx *c) { return c ? c->op.nP * c->op.aslm : 0; }
uint32_t tr_ctx_slm(const tr_ctx *c) { return c ? c->op.slm : 0; }
uint32_t tr_ctx_aslm(const tr_ctx *c) { return c ? c->op.aslm : 0; }
uint32_t tr_ctx_nP(const tr_ctx *c) { return c ? c->op.nP : 0; }

int tr_ctx_permutation(const tr_ctx *c, uint32_t *perm) {
    if (!c || !perm) return fail(TR_ERR_ARG, "tr_ctx_permutation: null argument");
    std::memcpy(perm, c->op.perm.data(), sizeof(uint32_t) * c->op.nP);
    return TR_OK;
}

void *tr_device_alloc(tr_ctx *c, size_t bytes) {
    if (!c) return nullptr;
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 1)) != hipSuccess) {
        fail(TR_ERR_HIP, "tr_device_alloc: hipMalloc failed");
        return nullptr;
    }
    if (hipMemsetAsync(p, 0, bytes, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        hipFree(p);
        fail(TR_ERR_HIP, "tr_device_alloc: memset failed");
        return nullptr;
    }
    c->raw.insert(p);
    return p;
}

double *tr_alloc_vector(tr_ctx *c) {
    if (!c) return nullptr;
    void *p = nullptr;
    size_t bytes = sizeof(double) * c->nint();
    if (hipMalloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) {
        fail(TR_ERR_HIP, "tr_alloc_vector: hipMalloc failed");
        return nullptr;
    }
    if (hipMemsetAsync(p, 0, bytes, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        hipFree(p);
        fail(TR_ERR_HIP, "tr_alloc_vector: memset failed");
        return nullptr;
    }
    c->vectors.insert(p);
    return (double *)p;
}

int tr_device_free(tr_ctx *c, void *mem) {
    if (!c) return fail(TR_ERR_ARG, "tr_device_free: null context");
    if (!mem) return TR_OK;
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (c->vectors.erase(mem) || c->raw.erase(mem)) {
        HIPCHECK(hipFree(mem));
        return TR_OK;
    }
    return fail(TR_ERR_ARG, "tr_device_free: pointer not allocated by this context");
}

int tr_copy_to_dev(tr_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c || !dst || !src) return fail(TR_ERR_ARG, "tr_copy_to_dev: null argument");
    if (c->vectors.count(dst)) {
        if (bytes != sizeof(double) * c->next())
            return fail(TR_ERR_ARG, "tr_copy_to_dev: vector copies must be N()*sizeof(REAL) bytes");
        HIPCHECK(hipMemcpyAsync(c->d_stage, src, bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHECK(launch_to_internal(c->d_stage, (double *)dst, c->d_perm, c->op.nP, c->op.slm,
                                    c->op.aslm, c->stream));
    } else {
        HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
    return TR_OK;
}

int tr_copy_to_host(tr_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c || !dst || !src) return fail(TR_ERR_ARG, "tr_copy_to_host: null argument");
    if (c->vectors.count(const_cast<void *>(src))) {
        if (bytes != sizeof(double) * c->next())
            return fail(TR_ERR_ARG, "tr_copy_to_host: vector copies must be N()*sizeof(REAL) bytes");
        HIPCHECK(launch_to_external((const double *)src, c->d_stage, c->d_iperm, c->op.nP,
                                    c->op.slm, c->op.aslm, c->stream));
        HIPCHECK(hipMemcpyAsync(dst, c->d_stage, bytes, hipMemcpyDeviceToHost, c->stream));
    } else {
        HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
    return TR_OK;
}

int tr_get_host_small(tr_ctx *c, const double *src, double *dst) {
    if (!c || !src || !dst) return fail(TR_ERR_ARG, "tr_get_host_small: null argument");
    std::vector<double> q(c->next());
    int rc = tr_copy_to_host(c, q.data(), src, sizeof(double) * q.size());
    if (rc) return rc;
    for (size_t i = 0; i < c->op.nP; i++)
        for (size_t j = 0; j < c->op.slm; j++) dst[i * c->op.slm + j] = q[i * c->op.aslm + j];
    return TR_OK;
}

int tr_compute_lhs(tr_ctx *c, const double *f, double *Af) {
    if (!c || !f || !Af) return fail(TR_ERR_ARG, "tr_compute_lhs: null argument");
    return lhs(c, f, Af);
}

int tr_compute_lhs_diag(tr_ctx *c, const double *f, double *Af) {
    if (!c || !f || !Af) return fail(TR_ERR_ARG, "tr_compute_lhs_diag: null argument");
    return lhs_diag(c, f, Af);
}

int tr_compute_rhs(tr_ctx *c, double *b) {
    if (!c || !b) return fail(TR_ERR_ARG, "tr_compute_rhs: null argument");
    HIPCHECK(launch_rhs(c->d_rhs, c->op.nP, c->op.slm, b, c->stream));
    return TR_OK;
}

namespace tr {
__global__ void k_scalar_guess(const float *__restrict__ y, const uint32_t *__restrict__ perm, uint32_t nP,
                               uint32_t slm, double scale, double *__restrict__ f) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)nP * slm;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t v = i / slm;
        f[i] = (i - v * slm) == 0 ? scale * (double)y[perm[v]] : 0.0;
    }
}
}  // namespace tr

// f = the guess f[v][00] = scale * y[v] (y device, caller's order), the
// learned initial guess of the solvers (with tr_solve_params.x0).
int tr_ctx_scalar_guess(tr_ctx *c, const float *y, double scale, double *f) {
    if (!c || !y || !f) return fail(TR_ERR_ARG, "tr_ctx_scalar_guess: null argument");
    hipLaunchKernelGGL(k_scalar_guess, dim3(1024), dim3(256), 0, c->stream, y, c->d_perm, c->op.nP, c->op.slm, scale,
                       f);
    HIPCHECK(hipGetLastError());
    return TR_OK;
}

// MeshData::ComputeMoments (MeshData.cpp:96-116) of a device vector: U[nP]
// and T[3i+j][nP] (device arrays) in the caller's vertex order.
int tr_compute_moments(tr_ctx *c, const double *f, double *U, double *T) {
    if (!c || !f || !U || !T) return fail(TR_ERR_ARG, "tr_compute_moments: null argument");
    HIPCHECK(launch_moments(f, c->d_perm, c->d_mom_val, c->d_mom_pos, c->op.nP, c->op.slm, U, T, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return TR_OK;
}

// MeshData::ComputeFlux (MeshData.cpp:118-136): W[j][nT] (device) from the
// device moments T of tr_compute_moments and the mesh the context was built
// from (its tets are uploaded for the call).
int tr_compute_flux(tr_ctx *c, const tr_mesh_view *mesh, const double *T, double *W) {
    if (!c || !mesh || !mesh->mesh || !T || !W) return fail(TR_ERR_ARG, "tr_compute_flux: null argument");
    if (mesh->nP != c->op.nP) return fail(TR_ERR_ARG, "tr_compute_flux: mesh does not match the context");
    tr_tetrahedron *d = nullptr;
    HIPCHECK(hipMalloc((void **)&d, sizeof(tr_tetrahedron) * std::max<uint32_t>(mesh->nT, 1)));
    hipError_t e = hipMemcpyAsync(d, mesh->mesh, sizeof(tr_tetrahedron) * mesh->nT, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = launch_flux(d, mesh->nT, mesh->nP, T, W, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(d);
    if (e != hipSuccess) return fail(TR_ERR_HIP, std::string("tr_compute_flux: ") + hipGetErrorString(e));
    return TR_OK;
}

// Outward flux density through each boundary face of the mesh view the
// context was built from, F[j] (device, the view's face order): the mean
// over the face of 4 pi sum_lm O_n[lm][00] f_lm.  At a vacuum face the
// even-parity form has psi- = sign(Omega.n) psi+, so the net outward flux
// is the integral of |Omega.n| psi+ over the sphere.  O_n is the one the
// operator's boundary term uses for the face (kernels.cu:414-425: the axis
// rule; with TR_CTX_GENERAL_NORMALS, faces off the axes take their own
// normal); reflecting faces carry none.  No reference function outputs it.
int tr_compute_boundary_flux(tr_ctx *c, const tr_mesh_view *mesh, const double *f, double *F) {
    if (!c || !mesh || !f || !F || (mesh->nF && !mesh->bnd))
        return fail(TR_ERR_ARG, "tr_compute_boundary_flux: null argument");
    if (mesh->nP != c->op.nP || !c->voff.empty())
        return fail(TR_ERR_ARG, "tr_compute_boundary_flux: mesh does not match the context (batches: per problem)");
    const uint32_t slm = c->op.slm, nF = mesh->nF;
    if (!nF) return TR_OK;
    int maxk = 0;
    while ((uint32_t)((1 + maxk) * (1 + 2 * maxk)) < slm) maxk++;
    std::vector<double> col((size_t)3 * slm);
    for (int n = 0; n < 3; n++)
        for (uint32_t lm = 0; lm < slm; lm++) col[(size_t)n * slm + lm] = c->op.onc[((size_t)n * slm + lm) * slm];
    std::vector<uint32_t> rec((size_t)4 * nF);
    for (uint32_t j = 0; j < nF; j++) {
        const tr_face &fc = mesh->bnd[j];
        for (int k = 0; k < 3; k++) {
            if (fc.p[k] >= mesh->nP) return fail(TR_ERR_ARG, "tr_compute_boundary_flux: face vertex out of range");
            rec[4 * j + k] = c->op.iperm[fc.p[k]];
        }
        uint32_t row;
        const double S = std::sqrt(fc.s[0] * fc.s[0] + fc.s[1] * fc.s[1] + fc.s[2] * fc.s[2]);
        const double mx = std::max({std::fabs(fc.s[0]), std::fabs(fc.s[1]), std::fabs(fc.s[2])});
        if (mesh->face_kind && mesh->face_kind[j] != 0) {
            row = ~0u;
        } else if (c->op.general_normals && S > 0 && mx < (1.0 - 1e-12) * S) {
            row = (uint32_t)(col.size() / slm);
            col.resize(col.size() + slm);
            abs_normal_column(maxk, fc.s, col.data() + (size_t)row * slm);
        } else {
            const double surf = std::fabs(fc.s[0] + fc.s[1] + fc.s[2]);
            row = (2 * std::fabs(fc.s[0]) > surf) ? 0 : ((2 * std::fabs(fc.s[1]) > surf) ? 1 : 2);
        }
        rec[4 * j + 3] = row;
    }
    void *d = nullptr;
    const size_t rb = sizeof(uint32_t) * rec.size(), cb = sizeof(double) * col.size();
    HIPCHECK(hipMalloc(&d, rb + cb));
    // synchronous copies: the host tables are temporaries
    hipError_t e = hipMemcpy(d, rec.data(), rb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy((char *)d + rb, col.data(), cb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_boundary_flux(f, d, (const double *)((char *)d + rb), nF, slm, F, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(d);
    if (e != hipSuccess) return fail(TR_ERR_HIP, std::string("tr_compute_boundary_flux: ") + hipGetErrorString(e));
    return TR_OK;
}

int tr_scale(tr_ctx *c, double *x, double wx) {
    if (!c || !x) return fail(TR_ERR_ARG, "tr_scale: null argument");
    return axpby(c, x, wx, nullptr, 0.0);
}
int tr_add_prod(tr_ctx *c, double *x, const double *y, double wy) {
    if (!c || !x || !y) return fail(TR_ERR_ARG, "tr_add_prod: null argument");
    return axpby(c, x, 1.0, y, wy);
}
int tr_mul_add(tr_ctx *c, double *x, double wx, const double *y) {
    if (!c || !x || !y) return fail(TR_ERR_ARG, "tr_mul_add: null argument");
    return axpby(c, x, wx, y, 1.0);
}
int tr_mul_add_prod(tr_ctx *c, double *x, double wx, const double *y, double wy) {
    if (!c || !x || !y) return fail(TR_ERR_ARG, "tr_mul_add_prod: null argument");
    return axpby(c, x, wx, y, wy);
}

double tr_norm(tr_ctx *c, const double *x) {
    if (!c || !x) {
        fail(TR_ERR_ARG, "tr_norm: null argument");
        return NAN;
    }
    return std::sqrt(dot(c, x, x));
}

double tr_dot(tr_ctx *c, const double *x, const double *y) {
    if (!c || !x || !y) {
        fail(TR_ERR_ARG, "tr_dot: null argument");
        return NAN;
    }
    return dot(c, x, y);
}

int tr_synchronize(tr_ctx *c) {
    if (!c) return fail(TR_ERR_ARG, "tr_synchronize: null context");
    HIPCHECK(hipStreamSynchronize(c->stream));
    return TR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Solver: main.cpp:125-162 (outer PCG, Fletcher-Reeves beta) with
// invprec (main.cpp:223-248, inner CG on the angular-diagonal operator).
namespace tr {

struct Work {
    tr_ctx *c;
    std::vector<double *> v;
    explicit Work(tr_ctx *c_) : c(c_) {}
    double *get() {
        double *p = nullptr;
        size_t bytes = sizeof(double) * c->nint();
        if (hipMalloc((void **)&p, std::max<size_t>(bytes, 8)) != hipSuccess) return nullptr;
        hipMemsetAsync(p, 0, bytes, c->stream);
        v.push_back(p);
        return p;
    }
    ~Work() {
        hipStreamSynchronize(c->stream);
        for (double *p : v) hipFree(p);
    }
};

static int invprec(tr_ctx *c, const tr_solve_params *prm, double *f, const double *b, double *Ap,
                   double *p, double *r, int *kout) {
    int rc;
    if ((rc = axpby(c, f, 0, b, 0)) || (rc = lhs_diag(c, f, Ap)) || (rc = axpby(c, r, 0, b, 1)) ||
        (rc = axpby(c, r, 1, Ap, -1)) || (rc = axpby(c, p, 0, r, 1)))
        return rc;
    int k = 0;
    double nrz = std::sqrt(dot(c, r, r));  // main.cpp:230-231: norm, then squared
    nrz *= nrz;
    while (k < prm->inner_maxit) {
        if ((rc = lhs_diag(c, p, Ap))) return rc;
        double alpha = nrz / dot(c, p, Ap);
        if ((rc = axpby(c, f, 1, p, alpha)) || (rc = axpby(c, r, 1, Ap, -alpha))) return rc;
        double nr = std::sqrt(dot(c, r, r));
        if (!(nr >= prm->inner_tol)) break;
        double nrz2 = nr * nr, beta = nrz2 / nrz;
        nrz = nrz2;
        if ((rc = axpby(c, p, beta, r, 1))) return rc;
        k++;
    }
    *kout = k;
    return TR_OK;
}

static int solve_reference(tr_ctx *c, const tr_solve_params *prm, double *f, tr_solve_stats *st,
                           double *hist) {
    Work w(c);
    double *b = w.get(), *p = w.get(), *Ap = w.get(), *z = w.get(), *r = w.get(), *pp = w.get(),
           *pAp = w.get(), *pr = w.get();
    if (!pr) return fail(TR_ERR_HIP, "tr_solve: out of device memory");
    int rc, preck = 0, ki;
    auto t0 = std::chrono::steady_clock::now();
    if ((rc = axpby(c, f, 0, nullptr, 0)) || (rc = tr_compute_rhs(c, b))) return rc;
    double bnorm = std::sqrt(dot(c, b, b));
    if ((rc = lhs(c, f, Ap)) || (rc = residual(c, r, b, Ap))) return rc;
    if (prm->use_prec) {
        if ((rc = invprec(c, prm, z, r, pAp, pp, pr, &ki))) return rc;
        preck += ki;
    } else if ((rc = axpby(c, z, 0, r, 1)))
        return rc;
    if ((rc = axpby(c, p, 0, z, 1))) return rc;
    int k = 0;
    double nrz = dot(c, r, z), nr = std::sqrt(dot(c, r, r));
    bool conv = false;
    while (k < prm->maxit) {
        if ((rc = lhs(c, p, Ap))) return rc;
        double alpha = nrz / dot(c, p, Ap);
        if ((rc = axpby(c, f, 1, p, alpha)) || (rc = axpby(c, r, 1, Ap, -alpha))) return rc;
        nr = std::sqrt(dot(c, r, r));
        conv = nr < prm->tol_abs || (prm->tol_rel > 0 && nr <= prm->tol_rel * bnorm);
        double beta = 0;
        if (!conv) {
            if (prm->use_prec) {
                if ((rc = invprec(c, prm, z, r, pAp, pp, pr, &ki))) return rc;
                preck += ki;
            } else if ((rc = axpby(c, z, 0, r, 1)))
                return rc;
            double nrz2 = dot(c, z, r);
            beta = nrz2 / nrz;
            nrz = nrz2;
            if ((rc = axpby(c, p, beta, z, 1))) return rc;
        }
        if (hist && k < prm->record) {
            hist[3 * k] = alpha;
            hist[3 * k + 1] = beta;
            hist[3 * k + 2] = nr;
        }
        if (conv || !std::isfinite(nr)) break;
        k++;
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
    auto t1 = std::chrono::steady_clock::now();
    // true residual with the FP64 operator
    if ((rc = lhs(c, f, Ap)) || (rc = residual(c, r, b, Ap))) return rc;
    double tr_ = std::sqrt(dot(c, r, r));
    if (st) {
        st->iters = k;
        st->preck = preck;
        st->rnorm = nr;
        st->bnorm = bnorm;
        st->true_rel = bnorm > 0 ? tr_ / bnorm : tr_;
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        st->status = conv ? TR_OK : TR_ERR_NOCONV;
    }
    return TR_OK;
}

}  // namespace tr

namespace tr {

// diag(A)[v][lm] from the assembled data: slot 0 of the volume stencil is
// the vertex itself; on the boundary the vertex's own boundary slot.
std::vector<double> operator_diagonal(const HostOperator &op) {
    const uint32_t nP = op.nP, slm = op.slm, NC = op.NC;
    std::vector<double> d((size_t)nP * slm);
    for (uint32_t v = 0; v < nP; v++)
        for (uint32_t lm = 0; lm < slm; lm++) {
            const double mfl = (!op.mf0.empty() && lm == 0) ? op.massfac[0] * op.mf0[v] : op.massfac[lm];
            double s = mfl * op.coef[(size_t)0 * nP + v];
            for (uint32_t c = 1; c < NC; c++)
                s += op.omc[((size_t)(c - 1) * slm + lm) * slm + lm] * op.coef[(size_t)c * nP + v];
            int b = op.bof[v];
            if (b >= 0)
                for (uint32_t q = 0; q < op.WB; q++)
                    if (op.bnbr[(size_t)q * op.nB + b] == v) {
                        for (int n = 0; n < 3; n++)
                            s += op.onc[((size_t)n * slm + lm) * slm + lm] *
                                 op.bcoef[((size_t)q * 3 + n) * op.nB + b];
                        break;
                    }
            d[(size_t)v * slm + lm] = s;
        }
    // faces off the axes: O(n_F)[lm][lm] (S_F/12) 2 for the vertex's own term
    for (size_t b = 0; b < op.gvert.size(); b++) {
        const uint32_t v = op.gvert[b];
        for (uint32_t j = op.gstart[b]; j < op.gstart[b + 1]; j++) {
            const uint32_t F = op.gface[j];
            for (uint32_t lm = 0; lm < slm; lm++)
                d[(size_t)v * slm + lm] += 2.0 * op.garea[F] * op.gOd[(size_t)F * slm + lm];
        }
    }
    return d;
}

// Starting point of the device solvers: 0, or the caller's f (prm->x0)
// scaled by alpha = (f.b)/(f.Af), the multiple of f closest to the solution
// in the energy norm: its squared A-norm error is that of 0 less
// (f.b)^2/(f.Af), so it is accepted whenever that scaling is defined
// (finite, f.Af > 0) and positive (f.b > 0; a guess anti-correlated with
// the source is rejected).  A rejected guess gives 0, reported as a
// fallback.  (A residual-norm test would reject good guesses: a scalar
// guess leaves the higher moments' rows of b - A x0 larger than b.)
static int start_point(tr_ctx *c, const tr_solve_params *prm, double *f, const double *b, double *r,
                       tr_solve_stats *st) {
    if (st) st->fallback = 0;
    if (prm->x0) {
        int rc = lhs(c, f, r);
        if (rc) return rc;
        const double fAf = dot(c, f, r), fb = dot(c, f, b);
        if (std::isfinite(fAf) && std::isfinite(fb) && fAf > 0 && fb > 0)
            return axpby(c, f, fb / fAf, nullptr, 0);
        if (st) st->fallback = 1;
    }
    // a fill, not 0 * f: a rejected guess may hold NaN or Inf
    HIPCHECK(hipMemsetAsync(f, 0, sizeof(double) * c->nint(), c->stream));
    return TR_OK;
}

// Chebyshev polynomial preconditioner (use_prec 4): z = the deg-th iterate
// of the Chebyshev iteration on D^-1 A z = D^-1 r from z = 0 over the
// interval [lambda_max / ratio, lambda_max] of D^-1 A's spectrum (Saad,
// Iterative Methods for Sparse Linear Systems, Alg. 12.1, in its three-term
// form).  Its residual polynomial is below 1 in magnitude on the interval
// and in (0, 1) below it, so M = p(D^-1 A) D^-1 is symmetric positive
// definite for any spectrum inside (0, lambda_max]: a fixed CG
// preconditioner.  lambda_max comes from an FP64 power iteration times 1.1.
// Each of the deg - 1 steps is one operator launch with the update fused
// into its write (matrix-core schedule).
int cheb_lambda(tr_ctx *c, const std::vector<double> &dinv) {
    if (c->cheb_lmax > 0) return TR_OK;
    const size_t n = c->nint();
    double *d = nullptr, *x = nullptr, *y = nullptr;
    HIPCHECK(hipMalloc((void **)&d, sizeof(double) * 3 * n));
    x = d + n;
    y = x + n;
    double lam = 0;
    int rc = TR_OK;
    if (hipMemcpyAsync(d, dinv.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = fail(TR_ERR_HIP, "chebyshev: upload failed");
    if (!rc) rc = lambda_max(c, d, x, y, lam);
    hipStreamSynchronize(c->stream);
    hipFree(d);
    if (rc) return rc;
    c->cheb_lmax = 1.1 * lam;
    return TR_OK;
}

static int solve_device(tr_ctx *c, const tr_solve_params *prm, double *f, tr_solve_stats *st) {
    Work w(c);
    const size_t n = c->nint();
    double *b = w.get(), *r = w.get(), *p = w.get(), *Ap = w.get();
    double *part = nullptr, *part2 = nullptr, *sc = nullptr, *dinv = nullptr;
    const uint32_t nb = apply_tiles(c->op.nP, c->map);
    HIPCHECK(hipMalloc((void **)&part, sizeof(double) * (std::max<size_t>(nb, 3 * 1024) + 2 * 1024 + 16)));
    w.v.push_back(part);
    part2 = part + std::max<size_t>(nb, 3 * 1024);
    sc = part2 + 2 * 1024;
    if (!Ap) return fail(TR_ERR_HIP, "tr_solve: out of device memory");
    int rc;
    ChebPrec<double> cheb;
    if (prm->use_prec == 5 && !prm->x0)
        return fail(TR_ERR_ARG, "tr_solve: use_prec 5 (learned coarse correction) takes the learned guess in f "
                                "with x0 = 1");
    // Under the FP64 K-blocked schedule the solve runs on vectors in its
    // harmonic order (the parity classes), so that no application permutes:
    // b and the start in, the solution out.  The multilevel and learned
    // preconditioners index harmonics in the internal order: they keep it.
    const uint16_t *hp = (prm->use_prec == 0 || prm->use_prec == 2 || prm->use_prec == 4) ? c->perm64() : nullptr;
    if (prm->use_prec == 2 || prm->use_prec == 4 || prm->use_prec == 5) {
        std::vector<double> d = operator_diagonal(c->op), dp;
        for (double &x : d) x = 1.0 / x;
        if (hp) {
            const uint32_t slm = c->op.slm;
            dp.resize(n);
            for (size_t v = 0; v < c->op.nP; v++)
                for (uint32_t k = 0; k < slm; k++) dp[v * slm + k] = d[v * slm + c->hp64[k]];
        }
        dinv = w.get();
        if (!dinv) return fail(TR_ERR_HIP, "tr_solve: out of device memory");
        HIPCHECK(hipMemcpyAsync(dinv, hp ? dp.data() : d.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                                c->stream));
        // d and dp are temporaries: the copy must have read them before they go
        HIPCHECK(hipStreamSynchronize(c->stream));
        if (prm->use_prec == 4) {
            // lambda_max by the power iteration on the internal order: d, not dp
            if ((rc = cheb_lambda(c, d))) return rc;
            cheb.a = hp ? c->args_cls() : c->args();
            cheb.NC = (int)c->op.NC;
            cheb.map = c->map;
            cheb.dinv = dinv;
            cheb.n = n;
            cheb.zb = w.get();
            cheb.tmp = w.get();
            if (!cheb.tmp) return fail(TR_ERR_HIP, "tr_solve: out of device memory");
            cheb.setup(c->cheb_deg, c->cheb_lmax, c->cheb_ratio);
        }
    } else if (prm->use_prec == 3) {
        if ((rc = build_multilevel(c, false, c->mg_coarsest, c->mg_agg))) return rc;
    } else if (prm->use_prec != 0) {
        return fail(TR_ERR_ARG, "tr_solve: TR_SOLVER_PCG_FP64 takes use_prec 0 (none), 2 (Jacobi), 3 (multilevel), "
                                "4 (Chebyshev) or 5 (Jacobi + learned coarse correction)");
    }
    tr_solve_stats st0{};
    if (!st) st = &st0;
    if ((rc = tr_compute_rhs(c, b)) || (rc = start_point(c, prm, f, b, r, st))) return rc;
    CgSpace<double> S;
    // the learned coarse correction: g = the scaled guess; without one (the
    // guess rejected) the solve is Jacobi-preconditioned
    double *gvec = nullptr, *ccp = nullptr;
    if (prm->use_prec == 5 && !st->fallback) {
        gvec = w.get();
        S.z = w.get();
        if (!gvec || !S.z) return fail(TR_ERR_HIP, "tr_solve: out of device memory");
        const int np = learned_coarse_partials();
        HIPCHECK(hipMalloc((void **)&ccp, sizeof(double) * (np + 8)));
        w.v.push_back(ccp);
        HIPCHECK(hipMemcpyAsync(gvec, f, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        if ((rc = lhs(c, gvec, Ap))) return rc;
        const double gAg = dot(c, gvec, Ap);
        if (!(gAg > 0)) return fail(TR_ERR_HIP, "tr_solve: learned coarse vector with g.Ag <= 0");
        HIPCHECK(hipMemcpy(ccp + np, &gAg, sizeof(double), hipMemcpyHostToDevice));
        const hipStream_t cs = c->stream;
        const double *gv = gvec, *dv = dinv;
        S.prec = [gv, dv, ccp, np, n, cs](const double *rr, double *z, const double *skip) {
            return launch_learned_coarse<double>(rr, z, dv, gv, n, ccp, ccp + np, skip, cs);
        };
    }
    if (prm->use_prec == 3) {
        S.z = w.get();
        if (!S.z) return fail(TR_ERR_HIP, "tr_solve: out of device memory");
        S.prec = [c](const double *rr, double *z, const double *skip) { return multilevel_apply64(c, rr, z, skip); };
    } else if (prm->use_prec == 4) {
        S.z = w.get();
        if (!S.z) return fail(TR_ERR_HIP, "tr_solve: out of device memory");
        const hipStream_t cs = c->stream;
        S.prec = [&cheb, cs](const double *rr, double *z, const double *skip) { return cheb.apply(rr, z, skip, cs); };
        S.prec_loop = [&cheb, cs](const double *rr, double *z, const double *skip, double *rzp) {
            return cheb.steps(rr, z, skip, cs, rzp);
        };
        S.pre_out = cheb.first_buffer(S.z);
        S.pre_dinv = dinv;
        S.pre_scale = cheb.inv_theta;
        S.prec_nparts = cheb.nparts();
    }
    S.a = hp ? c->args_cls() : c->args();
    S.NC = (int)c->op.NC;
    S.map = c->map;
    S.n = n;
    S.x = f;
    S.r = r;
    S.p = p;
    S.Ap = Ap;
    S.b = b;
    S.dinv = dinv;
    S.part = part;
    S.part2 = part2;
    S.sc = sc;
    if (hp) {
        // r and Ap are free until the solve: the start and b in the class order
        HIPCHECK(launch_perm_harm(f, r, hp, c->op.slm, n, false, nullptr, c->stream));
        HIPCHECK(launch_perm_harm(b, Ap, hp, c->op.slm, n, false, nullptr, c->stream));
        HIPCHECK(hipMemcpyAsync(f, r, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        HIPCHECK(hipMemcpyAsync(b, Ap, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    }
    CgParams cp;
    cp.maxit = prm->maxit;
    cp.tol_abs = prm->tol_abs;
    cp.tol_rel = prm->tol_rel;
    cp.chunk = 16;
    CgResult res;
    if ((rc = pcg_device<double>(S, cp, c->stream, res))) return rc;
    if (hp) {
        // the solution and b back in the internal order
        HIPCHECK(launch_perm_harm(f, r, hp, c->op.slm, n, true, nullptr, c->stream));
        HIPCHECK(launch_perm_harm(b, Ap, hp, c->op.slm, n, true, nullptr, c->stream));
        HIPCHECK(hipMemcpyAsync(f, r, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        HIPCHECK(hipMemcpyAsync(b, Ap, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    }
    if ((rc = lhs(c, f, Ap)) || (rc = residual(c, r, b, Ap))) return rc;
    double trn = std::sqrt(dot(c, r, r)), bn = std::sqrt(res.bb);
    if (st) {
        st->iters = res.iters;
        st->preck = 0;
        st->rnorm = std::sqrt(res.rr);
        st->bnorm = bn;
        st->true_rel = bn > 0 ? trn / bn : trn;
        st->seconds = res.seconds;
        st->status = res.flag == 1 ? TR_OK : TR_ERR_NOCONV;
    }
    return TR_OK;
}


static std::vector<float> to_f32(const std::vector<double> &v) {
    return std::vector<float>(v.begin(), v.end());
}

int ensure_fp32(tr_ctx *c) {
    if (c->have32) return TR_OK;
    const HostOperator &op = c->op;
    c->map32 = choose_map(op.slm, 8);
    if (c->ang_prec != 0 && v4_ok(op))
        c->map32 = mfma_map_ctx(c, 4, 2);
    else if (c->ang_prec != 0 && v5_ok(op))
        c->map32 = mfma5_map_ctx(c, 8);
    if (c->map32.variant <= 3) {
        c->map32.variant = valu_variant(op, 4, c->map32.variant);
        if (!fit_lds(c->map32, op, 4, c->device))
            return fail(TR_ERR_ARG, "tr_solve: angular order too large for LDS staging (FP32 operator)");
    }
    int rc;
    if ((rc = upload(c, &c->f_coef, to_f32(op.coef))) || (rc = upload(c, &c->f_bcoef, to_f32(op.bcoef))) ||
        (rc = upload(c, &c->f_aval, to_f32(op.aval))) || (rc = upload(c, &c->f_bval, to_f32(op.bval))) ||
        (rc = upload(c, &c->f_massfac, to_f32(op.massfac))) || (rc = upload(c, &c->f_ddiag, to_f32(op.ddiag))) ||
        (rc = upload(c, &c->f_bdiag, to_f32(op.bdiag))) || (rc = upload(c, &c->f_crec, coef_records<float>(op))) ||
        (rc = upload(c, &c->f_bcrec, bcoef_records<float>(op))) ||
        (!op.mf0.empty() && (rc = upload(c, &c->f_mf0, to_f32(op.mf0)))) ||
        (!op.gvert.empty() &&
         ((rc = upload(c, &c->f_garea, to_f32(op.garea))) || (rc = upload(c, &c->f_gR, to_f32(op.gR))) ||
          (rc = upload(c, &c->f_gOd, to_f32(op.gOd))) || (rc = upload(c, &c->f_gOz, to_f32(op.gOz))))) ||
        (rc = upload(c, (unsigned char **)&c->f_tab, table_records<float>(op))))
        return rc;
    if (v5_ok(op) && (rc = build_tables5(c))) return rc;
    // the persistent grids of variant 5 depend on its tables' LDS
    if (c->map32.variant == 5) c->map32 = mfma5_map_ctx(c, c->map32.WPB);
    if (mfma_supported(op.slm, 1) &&
        ((rc = upload(c, (unsigned char **)&c->f_tab4, mfma_tables(op, 1, 1))) ||
         (rc = upload(c, (unsigned char **)&c->f_tab4b1, mfma_tables(op, 2, 1))) ||
         (rc = upload(c, (unsigned char **)&c->f_tab4b3, mfma_tables(op, 2, 3)))))
        return rc;
    HIPCHECK(hipStreamSynchronize(c->stream));
    c->have32 = true;
    return TR_OK;
}

// FP64 iterative refinement around an FP32 device CG:
//   r = b - A64 x;  d = (A32)^-1 (r / |r|) to inner_tol;  x += |r| d
// By default every step after the first continues the previous step's CG
// instead of restarting it (residual replacement, cg_replace_residual): its
// residual becomes the new true FP64 residual and its last direction is
// kept, so the Krylov space built so far is not thrown away.
// The plain preconditioned FP64 solver (TR_SOLVER_PCG_FP64, same use_prec)
// takes over from the last good iterate when the FP64 residual fails to
// drop: a refinement step that does not cut it by kStagnation, an update
// that makes it larger or non-finite (taken back), or an inner breakdown.
// The inner CG stops by itself when its residual stalls
// (kInnerStallWindow).
// (Defaults; tr_ctx_set_refinement changes them per context.)
static constexpr double kStagnation = 0.5;
static constexpr int kInnerStallWindow = 100;

static int solve_mixed(tr_ctx *c, const tr_solve_params *prm, double *f, tr_solve_stats *st) {
    int rc = ensure_fp32(c);
    if (rc) return rc;
    // the multilevel hierarchy's coarse operators take the internal
    // harmonic order: under a reordered FP32 schedule this solve runs the
    // FP32 VALU schedule instead (restored on return)
    struct MapSwap {
        tr_ctx *c;
        LaneMap saved;
        bool on;
        ~MapSwap() {
            if (on) c->map32 = saved;
        }
    } swap{c, c->map32, false};
    if (prm->use_prec == 3 && c->perm32()) {
        LaneMap m = choose_map(c->op.slm, 8);
        m.variant = valu_variant(c->op, 4, m.variant);
        if (!fit_lds(m, c->op, 4, c->device))
            return fail(TR_ERR_ARG, "tr_solve: no FP32 schedule for the multilevel preconditioner");
        c->map32 = m;
        swap.on = true;
    }
    Work w(c);
    const size_t n = c->nint();
    double *b = w.get(), *r = w.get(), *Ap = w.get();
    float *v32 = nullptr;
    const uint32_t nb = apply_tiles(c->op.nP, c->map32);
    double *part = nullptr;
    HIPCHECK(hipMalloc((void **)&part, sizeof(double) * (std::max<size_t>(nb, 3 * 1024) + 2 * 1024 + 16)));
    w.v.push_back(part);
    HIPCHECK(hipMalloc((void **)&v32, sizeof(float) * (7 * n + 8)));
    w.v.push_back((double *)v32);
    float *d32 = v32, *r32 = v32 + n, *p32 = v32 + 2 * n, *Ap32 = v32 + 3 * n, *b32 = v32 + 4 * n;
    float *dinv32 = nullptr;
    if (!Ap) return fail(TR_ERR_HIP, "tr_solve: out of device memory");
    ChebPrec<float> cheb;
    if (prm->use_prec == 5 && !prm->x0)
        return fail(TR_ERR_ARG, "tr_solve: use_prec 5 (learned coarse correction) takes the learned guess in f "
                                "with x0 = 1");
    if (prm->use_prec == 2 || prm->use_prec == 4 || prm->use_prec == 5) {
        std::vector<double> d = operator_diagonal(c->op);
        std::vector<float> di(d.size());
        for (size_t i = 0; i < d.size(); i++) di[i] = (float)(1.0 / d[i]);
        if (c->perm32()) {  // into the FP32 vectors' harmonic order
            const uint32_t slm = c->op.slm;
            for (size_t v = 0; v < c->op.nP; v++)
                for (uint32_t k = 0; k < slm; k++) di[v * slm + k] = (float)(1.0 / d[v * slm + c->hp32[k]]);
        }
        dinv32 = v32 + 5 * n;
        HIPCHECK(hipMemcpyAsync(dinv32, di.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
        if (prm->use_prec == 4) {
            for (double &x : d) x = 1.0 / x;
            if ((rc = cheb_lambda(c, d))) return rc;
            float *ex = nullptr;
            HIPCHECK(hipMalloc((void **)&ex, sizeof(float) * 2 * n + 8));
            w.v.push_back((double *)ex);
            cheb.a = c->args32();
            // a lighter BF16 split for the steps of the 16-harmonic-tile
            // schedule: its couplings in BF16 alone are a fixed symmetric
            // perturbation, but the split of the gathered values rounds z
            // itself (about 2^-17 relative), so the polynomial is linear only
            // to that level.  Variant 5 keeps the operator's split: at 33..80
            // harmonics and ten channels the couplings' BF16 rounding
            // (relative to lambda_max) swamps the slow modes the polynomial
            // targets.
            if (c->cheb_split > 0 && c->cheb_split < cheb.a.nsplit && c->map32.variant == 4 &&
                c->map32.mfma_mode == 2)
                cheb.a.nsplit = c->cheb_split;
            cheb.NC = (int)c->op.NC;
            cheb.map = c->map32;
            cheb.dinv = dinv32;
            cheb.n = n;
            cheb.zb = ex;
            cheb.tmp = ex + n;
            cheb.setup(c->cheb_deg, c->cheb_lmax, c->cheb_ratio);
        }
    } else if (prm->use_prec == 3) {
        if ((rc = build_multilevel(c, true, c->mg_coarsest, c->mg_agg))) return rc;
    } else if (prm->use_prec != 0) {
        return fail(TR_ERR_ARG, "tr_solve: TR_SOLVER_MIXED takes use_prec 0 (none), 2 (Jacobi), 3 (multilevel), "
                                "4 (Chebyshev) or 5 (Jacobi + learned coarse correction)");
    }
    auto t0 = std::chrono::steady_clock::now();
    tr_solve_stats st0{};
    if (!st) st = &st0;
    if ((rc = tr_compute_rhs(c, b)) || (rc = start_point(c, prm, f, b, r, st))) return rc;
    const double bn = std::sqrt(dot(c, b, b));
    CgSpace<float> S;
    // the learned coarse correction of the inner solves: g = the scaled
    // guess in the FP32 vectors' order, g.Ag with the FP64 operator
    if (prm->use_prec == 5 && !st->fallback) {
        const int np = learned_coarse_partials();
        float *g32 = nullptr;
        double *ccp = nullptr;
        HIPCHECK(hipMalloc((void **)&g32, sizeof(float) * 2 * n + 16));
        w.v.push_back((double *)g32);
        HIPCHECK(hipMalloc((void **)&ccp, sizeof(double) * (np + 8)));
        w.v.push_back(ccp);
        S.z = g32 + n;
        if ((rc = lhs(c, f, Ap))) return rc;
        const double gAg = dot(c, f, Ap);
        if (!(gAg > 0)) return fail(TR_ERR_HIP, "tr_solve: learned coarse vector with g.Ag <= 0");
        HIPCHECK(launch_to_f32_scaled(f, g32, 1.0, n, c->stream, c->perm32(), c->op.slm));
        HIPCHECK(hipStreamSynchronize(c->stream));
        HIPCHECK(hipMemcpy(ccp + np, &gAg, sizeof(double), hipMemcpyHostToDevice));
        const hipStream_t cs = c->stream;
        const float *dv = dinv32;
        S.prec = [g32, dv, ccp, np, n, cs](const float *rr, float *z, const double *skip) {
            return launch_learned_coarse<float>(rr, z, dv, g32, n, ccp, ccp + np, skip, cs);
        };
    }
    if (prm->use_prec == 3) {
        S.z = v32 + 6 * n;
        S.prec = [c](const float *rr, float *z, const double *skip) { return multilevel_apply32(c, rr, z, skip); };
    } else if (prm->use_prec == 4) {
        S.z = v32 + 6 * n;
        const hipStream_t cs = c->stream;
        S.prec = [&cheb, cs](const float *rr, float *z, const double *skip) { return cheb.apply(rr, z, skip, cs); };
        S.prec_loop = [&cheb, cs](const float *rr, float *z, const double *skip, double *rzp) {
            return cheb.steps(rr, z, skip, cs, rzp);
        };
        S.pre_out = cheb.first_buffer(S.z);
        S.pre_dinv = dinv32;
        S.pre_scale = (float)cheb.inv_theta;
        S.prec_nparts = cheb.nparts();
    }
    S.a = c->args32();
    S.NC = (int)c->op.NC;
    S.map = c->map32;
    S.n = n;
    S.x = d32;
    S.r = r32;
    S.p = p32;
    S.Ap = Ap32;
    S.b = b32;
    S.dinv = dinv32;
    S.part = part;
    S.part2 = part + std::max<size_t>(nb, 3 * 1024);
    S.sc = S.part2 + 2 * 1024;
    CgParams cp;
    cp.maxit = prm->inner_maxit;
    cp.tol_abs = 0;
    cp.tol_rel = prm->inner_tol;
    cp.chunk = 16;
    cp.stall_window = c->ir_stall_window;
    // one iteration graph for every refinement step of this solve
    struct GraphKeep {
        hipGraphExec_t g = nullptr;
        ~GraphKeep() {
            if (g) hipGraphExecDestroy(g);
        }
    } keep_graph;
    cp.graph = &keep_graph.g;
    double *xb = w.get();  // the iterate before the last update
    if (!xb) return fail(TR_ERR_HIP, "tr_solve: out of device memory");
    int outer = 0, inner = 0, flag = TR_ERR_NOCONV;
    bool stalled = false, have_backup = false;
    double rn = bn, rn_prev = INFINITY;
    // residual replacement: the last inner solve stopped converged with a
    // direction, its scale and that direction's r.z
    bool resumable = false;
    double scale_prev = 0, rz_dir = 0, rr_rec = 0;
    for (; outer <= prm->maxit; outer++) {
        if ((rc = lhs(c, f, Ap)) || (rc = residual(c, r, b, Ap))) return rc;
        rn = std::sqrt(dot(c, r, r));
        // the last update made the true FP64 residual non-finite or larger:
        // take it back and finish with the FP64 solver
        if (have_backup && (!std::isfinite(rn) || rn > rn_prev)) {
            HIPCHECK(hipMemcpyAsync(f, xb, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
            rn = rn_prev;
            stalled = true;
            break;
        }
        if (!std::isfinite(rn)) break;
        if (rn <= prm->tol_rel * bn || rn < prm->tol_abs) {
            flag = TR_OK;
            break;
        }
        if (outer == prm->maxit) break;
        // a refinement step that did not cut the true residual by
        // kStagnation: the FP32 inner solve cannot make progress here
        if (have_backup && rn > c->ir_stagnation * rn_prev) {
            stalled = true;
            break;
        }
        rn_prev = rn;
        // the inner solve reduces its residual by inner_tol, or by what is
        // still needed (with a margin: the FP32 recursion overstates the
        // true reduction) when that is less
        {
            const double goal = std::max(prm->tol_rel * bn, prm->tol_abs);
            cp.tol_rel = std::max(prm->inner_tol, std::min(0.1, 0.6 * goal / rn));
        }
        CgResult res;
        bool resumed = false;
        HIPCHECK(hipMemsetAsync(d32, 0, sizeof(float) * n, c->stream));
        // continue the inner CG: its recursive residual replaced by the
        // true one in the new scale, its last direction kept; unless the
        // true residual (in the stopped solve's scale, rn / scale_prev) has
        // drifted from the recursive one by more than kReplaceDrift: then
        // the kept direction no longer fits the residual, and the step
        // restarts
        if (c->ir_reliable && resumable && rn / scale_prev <= kReplaceDrift * std::sqrt(rr_rec)) {
            HIPCHECK(launch_to_f32_scaled(r, r32, 1.0 / rn, n, c->stream, c->perm32(), c->op.slm));
            CgParams cq = cp;
            cq.resume = 1;
            if ((rc = cg_replace_residual<float>(S, cq, scale_prev / rn, rz_dir, c->stream, resumed))) return rc;
            if (resumed && (rc = pcg_device<float>(S, cq, c->stream, res))) return rc;
        }
        if (!resumed) {
            HIPCHECK(launch_to_f32_scaled(r, b32, 1.0 / rn, n, c->stream, c->perm32(), c->op.slm));
            if ((rc = pcg_device<float>(S, cp, c->stream, res))) return rc;
        }
        inner += res.iters;
        // a breakdown (NaN, or p.Ap <= 0 from an operator that is not SPD in
        // FP32) leaves no usable correction: the update is not applied
        if (res.flag == 3) {
            stalled = true;
            break;
        }
        resumable = res.flag == 1 && res.iters > 0;
        scale_prev = rn;
        rz_dir = res.rz_dir;
        rr_rec = res.rr;
        HIPCHECK(hipMemcpyAsync(xb, f, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        have_backup = true;
        HIPCHECK(launch_add_f32_scaled(f, d32, rn, n, c->stream, c->perm32(), c->op.slm));
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
    int fb_iters = 0;
    if (stalled) {
        // the plain preconditioned FP64 solver from the last good iterate
        tr_solve_params q = *prm;
        q.solver = TR_SOLVER_PCG_FP64;
        q.x0 = 1;
        // the caller's budget: inner_maxit per refinement step left
        q.maxit = std::max(prm->inner_maxit, 1) * std::max(1, prm->maxit - outer + 1);
        tr_solve_stats s2{};
        if ((rc = solve_device(c, &q, f, &s2))) return rc;
        fb_iters = s2.iters;
        rn = s2.true_rel * bn;
        flag = s2.status;
    }
    auto t1 = std::chrono::steady_clock::now();
    if (st) {
        const int fb0 = st->fallback;
        st->iters = outer + fb_iters;
        st->preck = inner;
        st->rnorm = rn;
        st->bnorm = bn;
        st->true_rel = bn > 0 ? rn / bn : rn;
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        st->status = flag;
        st->fallback = stalled ? 2 : fb0;
    }
    return TR_OK;
}
}  // namespace tr

extern "C" {

int tr_solve(tr_ctx *c, const tr_solve_params *prm, double *f, tr_solve_stats *st, double *hist) {
    if (!c || !prm || !f) return fail(TR_ERR_ARG, "tr_solve: null argument");
    if (!c->vectors.count(f)) return fail(TR_ERR_ARG, "tr_solve: f must come from tr_alloc_vector");
    switch (prm->solver) {
        case TR_SOLVER_REFERENCE:
            return solve_reference(c, prm, f, st, hist);
        case TR_SOLVER_PCG_FP64:
            return solve_device(c, prm, f, st);
        case TR_SOLVER_MIXED:
            return solve_mixed(c, prm, f, st);
        default:
            return fail(TR_ERR_ARG, "tr_solve: unknown solver");
    }
}

double tr_bench_lhs(tr_ctx *c, int reps) {
    if (!c || reps < 1) return NAN;
    Work w(c);
    double *f = w.get(), *r = w.get();
    if (!r) return NAN;
    std::vector<float> ms(reps);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    lhs(c, f, r);
    for (int i = 0; i < reps; i++) {
        hipEventRecord(e0, c->stream);
        lhs(c, f, r);
        hipEventRecord(e1, c->stream);
        hipEventSynchronize(e1);
        hipEventElapsedTime(&ms[i], e0, e1);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    std::sort(ms.begin(), ms.end());
    return ms[reps / 2] * 1e-3;
}

}  // extern "C"

namespace tr {
__global__ void k_copy_calib(const double4 *__restrict__ src, double4 *__restrict__ dst, size_t n);
}

extern "C" int tr_bench_operator(tr_ctx *c, int precision, int warmup, int steps, double *out) {
    if (!c || !out || steps < 1 || warmup < 0 || precision < 0 || precision > 4)
        return fail(TR_ERR_ARG, "tr_bench_operator: bad argument");
    int rc;
    // 3: the FP32 operator with its fused f.Af dot (the launch of a solver
    // iteration); 4: the same after a 1 GiB copy before each application,
    // so the caches hold what a solve's vector passes leave in them
    const bool dotp = precision >= 3, flush = precision == 4;
    if (dotp) precision = 1;
    if (precision == 1 && (rc = ensure_fp32(c))) return rc;
    Work w(c);
    double *f = w.get(), *r = w.get(), *part = dotp ? w.get() : nullptr;
    if (!r || (dotp && !part)) return fail(TR_ERR_HIP, "tr_bench_operator: out of device memory");
    const size_t fl_n = (size_t)1 << 25;  // 32-byte elements: 1 GiB read and written
    double4 *fl_a = nullptr, *fl_b = nullptr;
    if (flush) {
        if (hipMalloc((void **)&fl_a, fl_n * 32) != hipSuccess) return fail(TR_ERR_HIP, "tr_bench_operator: out of device memory");
        if (hipMalloc((void **)&fl_b, fl_n * 32) != hipSuccess) {
            hipFree(fl_a);
            return fail(TR_ERR_HIP, "tr_bench_operator: out of device memory");
        }
        hipMemsetAsync(fl_a, 0, fl_n * 32, c->stream);
    }
    struct FlushFree {
        double4 *a, *b;
        hipStream_t s;
        ~FlushFree() {
            if (a) {
                hipStreamSynchronize(s);
                hipFree(a);
                hipFree(b);
            }
        }
    } fl_free{fl_a, fl_b, c->stream};
    // a non-trivial input: f = rhs (converted to FP32 for precision 1, in
    // the first half of the FP64 buffer f; the result goes to r)
    HIPCHECK(launch_rhs(c->d_rhs, c->op.nP, c->op.slm, f, c->stream));
    float *f32 = reinterpret_cast<float *>(r), *r32 = f32 + c->nint();
    if (precision == 1) HIPCHECK(launch_to_f32_scaled(f, f32, 1.0, c->nint(), c->stream, c->perm32(), c->op.slm));
    auto apply = [&]() -> hipError_t {
        if (precision == 1)
            return launch_apply<float>(c->args32(), (int)c->op.NC, c->map32, f32, r32, part, nullptr, c->stream);
        // precision 2: the FP64 operator on vectors in its own harmonic
        // order (the K-blocked schedule's kernel alone, as its solves run it)
        return launch_apply<double>(precision == 2 ? c->args_cls() : c->args(), (int)c->op.NC, c->map, f, r,
                                    nullptr, nullptr, c->stream);
    };
    for (int i = 0; i < warmup; i++) HIPCHECK(apply());
    std::vector<hipEvent_t> ev(2 * (size_t)steps);
    for (auto &e : ev) HIPCHECK(hipEventCreate(&e));
    HIPCHECK(hipStreamSynchronize(c->stream));
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < steps; i++) {
        if (flush) hipLaunchKernelGGL(k_copy_calib, dim3(256 * 8 * 4), dim3(256), 0, c->stream, fl_a, fl_b, fl_n);
        HIPCHECK(hipEventRecord(ev[2 * i], c->stream));
        HIPCHECK(apply());
        HIPCHECK(hipEventRecord(ev[2 * i + 1], c->stream));
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
    auto t1 = std::chrono::steady_clock::now();
    double sum = 0, mn = 1e30, mx = 0;
    for (int i = 0; i < steps; i++) {
        float a = 0;
        HIPCHECK(hipEventElapsedTime(&a, ev[2 * i], ev[2 * i + 1]));
        sum += a;
        mn = std::min(mn, (double)a);
        mx = std::max(mx, (double)a);
    }
    for (auto &e : ev) hipEventDestroy(e);
    out[0] = std::chrono::duration<double>(t1 - t0).count();
    out[1] = sum / steps;
    out[2] = mn;
    out[3] = mx;
    return TR_OK;
}

// HIP kernel name (as a profiler prints it) of the operator schedule in use.
extern "C" int tr_ctx_kernel_name(tr_ctx *c, int precision, char *buf, size_t cap) {
    if (!c || !buf || cap < 2 || precision < 0 || precision > 1) return fail(TR_ERR_ARG, "tr_ctx_kernel_name: bad argument");
    if (precision == 1) {
        int rc = ensure_fp32(c);
        if (rc) return rc;
    }
    const LaneMap &m = precision ? c->map32 : c->map;
    const char *T = precision ? "float" : "double";
    const int NC = (int)c->op.NC, W = (int)c->op.W;
    std::string name;
    char tmp[160];
    if (m.variant == 5 && precision == 0)
        snprintf(tmp, sizeof tmp, "k_apply5d<%d, %d, %d, %d, 0>", NC, W, (int)c->b5d.nb, NC == 10 ? 5 : 7);
    else if (m.variant == 5 && c->b5.stat)
        snprintf(tmp, sizeof tmp, "k_apply5s<%d, %d, %d, %d, 0>", NC, W, (int)c->b5.stat, c->args32(m).nsplit);
    else if (m.variant == 5)
        snprintf(tmp, sizeof tmp, "k_apply5<%d, %d, %d, 0>", NC, W, (int)c->b5.nb);
    else if (m.variant == 4)
        snprintf(tmp, sizeof tmp, "k_apply4<%s, %d, %d, %d, %d, 0>", T, precision ? m.mfma_mode : 0, m.J, NC, W);
    else if (m.variant == 3 && (W == 15 || W == 16))
        snprintf(tmp, sizeof tmp, "k_apply3<%s, %d, %d, %d, false>", T, m.J, NC, W);
    else if (m.variant >= 2 && W <= TR_WMAX)
        snprintf(tmp, sizeof tmp, "k_apply2<%s, %d, %d, false>", T, m.J, NC);
    else
        snprintf(tmp, sizeof tmp, "k_apply<%s, %d, %d, false>", T, m.J, NC);
    name = tmp;
    std::strncpy(buf, name.c_str(), cap - 1);
    buf[cap - 1] = 0;
    return TR_OK;
}

namespace tr {
__global__ void k_copy_calib(const double4 *__restrict__ src, double4 *__restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
}  // namespace tr

// A plain device copy of `bytes` (rounded down to 32 B), timed with HIP
// events: the calibration kernel for memory-traffic counters and the
// achievable-bandwidth reference.  out[0] mean ms per copy.
extern "C" int tr_bench_copy(tr_ctx *c, size_t bytes, int reps, double *out) {
    if (!c || !out || reps < 1 || bytes < 32) return fail(TR_ERR_ARG, "tr_bench_copy: bad argument");
    size_t n = bytes / 32;
    double4 *a = nullptr, *b = nullptr;
    HIPCHECK(hipMalloc((void **)&a, n * 32));
    if (hipMalloc((void **)&b, n * 32) != hipSuccess) {
        hipFree(a);
        return fail(TR_ERR_HIP, "tr_bench_copy: hipMalloc failed");
    }
    hipMemsetAsync(a, 0, n * 32, c->stream);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    unsigned g = 256 * 8 * 4;
    hipLaunchKernelGGL(k_copy_calib, dim3(g), dim3(256), 0, c->stream, a, b, n);
    hipEventRecord(e0, c->stream);
    for (int i = 0; i < reps; i++) hipLaunchKernelGGL(k_copy_calib, dim3(g), dim3(256), 0, c->stream, a, b, n);
    hipEventRecord(e1, c->stream);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    if (e != hipSuccess) return fail(TR_ERR_HIP, hipGetErrorString(e));
    out[0] = ms / reps;
    return TR_OK;
}

namespace tr {

// Times every candidate schedule of the FP64 (precision 0) or FP32
// (precision 1) operator on this context's mesh and keeps the fastest.
static int autotune(tr_ctx *c, int precision, std::string &report) {
    const HostOperator &op = c->op;
    const bool f32 = precision == 1;
    int rc;
    if (f32 && (rc = ensure_fp32(c))) return rc;
    int maxlds = 65536;
    hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
    const size_t elem = f32 ? 4 : 8, n = c->nint();
    void *x = nullptr, *y = nullptr;
    HIPCHECK(hipMalloc(&x, n * elem));
    if (hipMalloc(&y, n * elem) != hipSuccess) {
        hipFree(x);
        return fail(TR_ERR_HIP, "autotune: out of device memory");
    }
    hipMemsetAsync(x, 0, n * elem, c->stream);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    LaneMap best = f32 ? c->map32 : c->map;
    float bestms = 1e30f;
    char line[160];
    // Candidates of equal accuracy only: a BF16 angular projection (FP32
    // operator with ang_prec 1 or 3) is the matrix-core schedule alone; the
    // exact schedules (FP64, or FP32 with ang_prec 0) race the VALU variants
    // against the FP64/FP32 MFMA one.
    std::vector<LaneMap> cands;
    const bool bf16 = f32 && c->ang_prec != 0 && (v4_ok(op) || v5_ok(op));
    if (!bf16)
        for (LaneMap m : map_candidates(op.slm, 11))
            for (int variant : {1, 2, 3})
                for (int wpb : {1, 2, 4}) {
                    if (variant >= 2 && op.W > 16) continue;
                    if (variant == 3 && !v3_ok(op, elem)) continue;
                    m.variant = variant;
                    m.WPB = wpb;
                    cands.push_back(m);
                }
    if (v4_ok(op))
        for (int wpb : {1, 2, 4, 8}) {
            LaneMap m = mfma_map_ctx(c, wpb, f32 ? (bf16 ? 2 : 1) : 0);
            cands.push_back(m);
        }
    if (bf16 && v5_ok(op))
        for (int wpb : {2, 4, 6, 8}) cands.push_back(mfma5_map_ctx(c, wpb));
    if (!f32 && c->d_tab5d) {
        // one 512-lane workgroup per CU (the FP64 blocks' LDS): WPB 8 only
        LaneMap m;
        if ((rc = mfma5d_map_ctx(c, 8, m))) {
            hipFree(x);
            hipFree(y);
            return rc;
        }
        cands.push_back(m);
    }
    {
        for (LaneMap m : cands) {
            {
                const int variant = m.variant, wpb = m.WPB;
                (void)variant;
                (void)wpb;
                if (lds_need(m, op, elem, c->ang_prec == 1 ? 1 : 3) > (size_t)maxlds) continue;
                auto run = [&]() {
                    if (f32) {
                        OpArgs<float> a = c->args32(m);
                        return launch_apply<float>(a, (int)op.NC, m, (const float *)x, (float *)y, nullptr,
                                                   nullptr, c->stream);
                    }
                    OpArgs<double> a = c->args(m);
                    return launch_apply<double>(a, (int)op.NC, m, (const double *)x, (double *)y, nullptr,
                                                nullptr, c->stream);
                };
                if (run() != hipSuccess) continue;
                hipEventRecord(e0, c->stream);
                const int reps = 5;
                for (int i = 0; i < reps; i++) run();
                hipEventRecord(e1, c->stream);
                if (hipEventSynchronize(e1) != hipSuccess) continue;
                float ms = 0;
                hipEventElapsedTime(&ms, e0, e1);
                ms /= reps;
                snprintf(line, sizeof line, "%s J=%d L=%d VPW=%d WPB=%d variant=%d mfma_mode=%d: %.2f us\n",
                         f32 ? "fp32" : "fp64", m.J, m.L, m.VPW, m.WPB, m.variant, m.variant == 4 ? m.mfma_mode : 0,
                         ms * 1e3);
                report += line;
                if (ms < bestms) {
                    bestms = ms;
                    best = m;
                }
            }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(x);
    hipFree(y);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(TR_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
    snprintf(line, sizeof line, "chosen %s J=%d WPB=%d variant=%d (%.2f us)\n", f32 ? "fp32" : "fp64", best.J,
             best.WPB, best.variant, bestms * 1e3);
    report += line;
    if (f32)
        c->map32 = best;
    else
        c->map = best;
    return TR_OK;
}

}  // namespace tr

extern "C" int tr_ctx_autotune(tr_ctx *c, int precision, char *report, size_t cap) {
    if (!c || precision < 0 || precision > 1) return fail(TR_ERR_ARG, "tr_ctx_autotune: bad argument");
    std::string rep;
    int rc = autotune(c, precision, rep);
    if (report && cap) {
        std::strncpy(report, rep.c_str(), cap - 1);
        report[cap - 1] = 0;
    }
    return rc;
}

extern "C" int tr_ctx_schedule(tr_ctx *c, int precision, int *out5) {
    if (!c || !out5 || precision < 0 || precision > 1) return fail(TR_ERR_ARG, "tr_ctx_schedule: bad argument");
    if (precision == 1) {
        int rc = ensure_fp32(c);
        if (rc) return rc;
    }
    const LaneMap &m = precision ? c->map32 : c->map;
    out5[0] = m.J;
    out5[1] = m.L;
    out5[2] = m.VPW;
    out5[3] = m.WPB;
    out5[4] = m.variant;
    return TR_OK;
}

extern "C" int tr_ctx_set_schedule(tr_ctx *c, int precision, int J, int WPB, int variant) {
    if (!c || precision < 0 || precision > 1) return fail(TR_ERR_ARG, "tr_ctx_set_schedule: bad argument");
    if (precision == 1) {
        int rc = ensure_fp32(c);
        if (rc) return rc;
    }
    const HostOperator &op = c->op;
    if (variant == 5) {
        if (precision == 0) {
            if (!c->d_tab5d || WPB != 8)
                return fail(TR_ERR_ARG, "tr_ctx_set_schedule: FP64 K-blocked matrix-core schedule unsupported here");
            LaneMap m;
            int rc = mfma5d_map_ctx(c, WPB, m);
            if (rc) return rc;
            c->map = m;
            return TR_OK;
        }
        if (c->ang_prec == 0 || !v5_ok(op) || WPB < 1 || WPB > 8)
            return fail(TR_ERR_ARG, "tr_ctx_set_schedule: K-blocked matrix-core schedule unsupported here");
        // J = 1: the masked block layout even where the static class layout
        // applies; otherwise the static one where it applies
        const bool want_static = J != 1;
        if (want_static != c->tab5_static) {
            c->tab5_static = want_static;
            int rc = build_tables5(c);
            if (rc) return rc;
        }
        c->map32 = mfma5_map_ctx(c, WPB);
        return TR_OK;
    }
    if (variant == 4) {
        if (!v4_ok(op) || (WPB != 1 && WPB != 2 && WPB != 4 && WPB != 8))
            return fail(TR_ERR_ARG, "tr_ctx_set_schedule: matrix-core schedule unsupported here");
        LaneMap m = mfma_map_ctx(c, WPB, precision ? (c->ang_prec == 0 ? 1 : 2) : 0);
        int maxlds = 65536;
        hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
        if (lds_need(m, op, precision ? 4 : 8, c->ang_prec == 1 ? 1 : 3) > (size_t)maxlds)
            return fail(TR_ERR_ARG, "tr_ctx_set_schedule: LDS budget exceeded");
        if (precision)
            c->map32 = m;
        else
            c->map = m;
        return TR_OK;
    }
    if (J < 1 || op.slm % J || (int)(op.slm / J) > 64 || (WPB != 1 && WPB != 2 && WPB != 4) ||
        (variant < 1 || variant > 3) || (variant >= 2 && op.W > 16) ||
        (variant == 3 && !v3_ok(op, precision ? 4 : 8)))
        return fail(TR_ERR_ARG, "tr_ctx_set_schedule: unsupported schedule");
    bool ok = false;
    for (const LaneMap &m : map_candidates(op.slm, 11)) ok = ok || m.J == J;
    if (!ok) return fail(TR_ERR_ARG, "tr_ctx_set_schedule: J not instantiated for this slm");
    LaneMap m;
    m.J = J;
    m.L = (int)(op.slm / J);
    m.VPW = 64 / m.L;
    m.WPB = WPB;
    m.variant = variant;
    int maxlds = 65536;
    hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
    if (lds_need(m, op, precision ? 4 : 8) > (size_t)maxlds)
        return fail(TR_ERR_ARG, "tr_ctx_set_schedule: LDS budget exceeded");
    if (precision)
        c->map32 = m;
    else
        c->map = m;
    return TR_OK;
}

// Af = A f with the FP32 operator (precision 1) or FP64 (precision 0); f and
// Af are FP64 device vectors, converted around the FP32 application.
extern "C" int tr_compute_lhs_prec(tr_ctx *c, int precision, const double *f, double *Af) {
    if (!c || !f || !Af || precision < 0 || precision > 1) return fail(TR_ERR_ARG, "tr_compute_lhs_prec: bad argument");
    if (precision == 0) return lhs(c, f, Af);
    int rc = ensure_fp32(c);
    if (rc) return rc;
    const size_t n = c->nint();
    float *v = nullptr;
    HIPCHECK(hipMalloc((void **)&v, sizeof(float) * 2 * n + 16));
    hipError_t e = launch_to_f32_scaled(f, v, 1.0, n, c->stream, c->perm32(), c->op.slm);
    if (e == hipSuccess)
        e = launch_apply<float>(c->args32(), (int)c->op.NC, c->map32, v, v + n, nullptr, nullptr, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(Af, 0, sizeof(double) * n, c->stream);
    if (e == hipSuccess) e = launch_add_f32_scaled(Af, v + n, 1.0, n, c->stream, c->perm32(), c->op.slm);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(v);
    if (e != hipSuccess) return fail(TR_ERR_HIP, std::string("tr_compute_lhs_prec: ") + hipGetErrorString(e));
    return TR_OK;
}

namespace tr {
// x[i] *= s[i] over the internal [v][slm] entries
__global__ __launch_bounds__(256) void k_mul_elem(double *__restrict__ x, const double *__restrict__ s, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] *= s[i];
}

// Extreme eigenvalues of the symmetric tridiagonal matrix (diagonal a[m],
// off-diagonal b[m-1]) by Sturm-sequence bisection: the number of
// eigenvalues below x is the number of negative pivots of T - x I.
static void tridiag_extremes(const std::vector<double> &a, const std::vector<double> &b, double &lo, double &hi) {
    const size_t m = a.size();
    double gl = 1e300, gh = -1e300;
    for (size_t i = 0; i < m; i++) {
        const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
        gl = std::min(gl, a[i] - r);
        gh = std::max(gh, a[i] + r);
    }
    auto below = [&](double x) {
        int cnt = 0;
        double d = 1.0;
        for (size_t i = 0; i < m; i++) {
            d = (a[i] - x) - (i > 0 ? b[i - 1] * b[i - 1] / d : 0.0);
            if (d == 0.0) d = -1e-300;
            if (d < 0) cnt++;
        }
        return cnt;
    };
    auto kth = [&](int k) {  // k-th smallest (0-based)
        double l = gl, h = gh;
        for (int it = 0; it < 200 && h - l > 1e-15 * std::max(std::fabs(l), std::fabs(h)); it++) {
            const double x = 0.5 * (l + h);
            (below(x) > k ? h : l) = x;
        }
        return 0.5 * (l + h);
    };
    lo = kth(0);
    hi = kth((int)m - 1);
}
}  // namespace tr

// Lanczos estimate of the extreme eigenvalues (Eigenvalue.cpp:53-88, eigest):
// `steps` Lanczos steps on A (jacobi = 0) or on D^-1/2 A D^-1/2, whose
// spectrum is that of D^-1 A (jacobi = 1), from the right-hand side b as
// the start vector (main.cpp:117-123), stopping early when the largest Ritz
// value changes by less than tol of itself (eigest: 1e-6; tol = 0 runs all
// steps, for the smallest Ritz value, which converges more slowly).  out[0] = smallest Ritz value,
// out[1] = largest, out[2] = the largest extrapolated as eigest does
// (lmax + (lmax - lold) / (1 - q), q = (lmax - lold) / (lold - loold)), out[3]
// = the number of steps taken.  The Ritz values are the tridiagonal's
// extreme eigenvalues by Sturm bisection, with the off-diagonal beta_{j+1}
// coupling rows j and j+1 (the reference's trieig is handed b with b[0] = 0
// in front and so couples each pair of rows through the previous beta).
int tr_eig_estimate(tr_ctx *c, int steps, int jacobi, double tol, double *out4) {
    if (!c || !out4 || steps < 1 || steps > 1000 || !(tol >= 0)) return fail(TR_ERR_ARG, "tr_eig_estimate: bad argument");
    const size_t n = c->nint();
    Work w(c);
    double *v = w.get(), *v0 = w.get(), *wv = w.get(), *tmp = w.get(), *sc = nullptr;
    if (!tmp) return fail(TR_ERR_HIP, "tr_eig_estimate: out of device memory");
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    if (jacobi) {
        std::vector<double> d = operator_diagonal(c->op);
        for (double &x : d) x = 1.0 / std::sqrt(x);
        sc = w.get();
        if (!sc) return fail(TR_ERR_HIP, "tr_eig_estimate: out of device memory");
        HIPCHECK(hipMemcpyAsync(sc, d.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
    }
    int rc = tr_compute_rhs(c, v);
    if (rc) return rc;
    const double nv = std::sqrt(dot(c, v, v));
    if (!(nv > 0)) return fail(TR_ERR_ARG, "tr_eig_estimate: zero right-hand side");
    if ((rc = axpby(c, v, 1.0 / nv, nullptr, 0.0)) || (rc = axpby(c, v0, 0.0, nullptr, 0.0))) return rc;
    std::vector<double> a, b;
    double beta = 0, lmax = 0, lold = 0, loold = 0, lmin = 0;
    int j = 0;
    for (; j < steps; j++) {
        // wv = S A S v, S = D^-1/2 (jacobi) or I
        if (sc) {
            HIPCHECK(hipMemcpyAsync(tmp, v, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
            hipLaunchKernelGGL(k_mul_elem, dim3(g), dim3(256), 0, c->stream, tmp, sc, n);
            if ((rc = lhs(c, tmp, wv))) return rc;
            hipLaunchKernelGGL(k_mul_elem, dim3(g), dim3(256), 0, c->stream, wv, sc, n);
        } else if ((rc = lhs(c, v, wv))) {
            return rc;
        }
        if ((rc = axpby(c, wv, 1.0, v0, -beta))) return rc;
        const double alpha = dot(c, wv, v);
        if ((rc = axpby(c, wv, 1.0, v, -alpha))) return rc;
        a.push_back(alpha);
        beta = std::sqrt(dot(c, wv, wv));
        loold = lold;
        lold = lmax;
        tridiag_extremes(a, b, lmin, lmax);
        if ((j > 0 && std::fabs(lmax - lold) < tol * std::fabs(lmax)) || !(beta > 0)) {
            j++;
            break;
        }
        b.push_back(beta);
        std::swap(v0, v);  // v0 = v_j, v = (w - alpha v_j - beta v_{j-1}) / beta_{j+1}
        if ((rc = axpby(c, v, 0.0, wv, 1.0 / beta))) return rc;
    }
    const double q = (lold != loold) ? (lmax - lold) / (lold - loold) : 0.0;
    out4[0] = lmin;
    out4[1] = lmax;
    out4[2] = (q != 1.0) ? lmax + (lmax - lold) / (1.0 - q) : lmax;
    out4[3] = j;
    return TR_OK;
}

// Chebyshev polynomial preconditioner (use_prec 4): its degree (1 = a
// scaled Jacobi step) and the ratio lambda_max / lambda_min of the
// interval of D^-1 A's spectrum it targets.
int tr_ctx_set_chebyshev(tr_ctx *c, int degree, double ratio, int step_split) {
    if (!c || degree < 1 || degree > 64 || !(ratio > 1.0) || step_split < 0 || step_split > 3)
        return fail(TR_ERR_ARG, "tr_ctx_set_chebyshev: bad argument");
    c->cheb_deg = degree;
    c->cheb_ratio = ratio;
    c->cheb_split = step_split;
    return TR_OK;
}

// Shape of the multilevel preconditioner (use_prec 3): levels are coarsened
// until at most `coarsest` vertices remain, by aggregates of `agg` vertices,
// and the coarsest level takes `nu` damped-Jacobi steps.  Drops a hierarchy
// already built.
extern "C" int tr_ctx_set_multilevel(tr_ctx *c, uint32_t coarsest, uint32_t agg, int nu) {
    if (!c || agg < 2 || nu < 1 || nu > 64) return fail(TR_ERR_ARG, "tr_ctx_set_multilevel: bad argument");
    c->mg_coarsest = coarsest;
    c->mg_agg = agg;
    c->mg_nu = nu;
    if (c->mg) {
        hipStreamSynchronize(c->stream);
        destroy_multilevel(c->mg);
        c->mg = nullptr;
    }
    return TR_OK;
}

extern "C" int tr_ctx_set_refinement(tr_ctx *c, double stagnation, int inner_stall_window) {
    if (!c || !(stagnation > 0) || inner_stall_window < 0)
        return fail(TR_ERR_ARG, "tr_ctx_set_refinement: stagnation must be > 0 and the window >= 0");
    c->ir_stagnation = stagnation;
    c->ir_stall_window = inner_stall_window;
    return TR_OK;
}

extern "C" int tr_ctx_set_residual_replacement(tr_ctx *c, int on) {
    if (!c || (on != 0 && on != 1)) return fail(TR_ERR_ARG, "tr_ctx_set_residual_replacement: 0 or 1");
    c->ir_reliable = on;
    return TR_OK;
}

// Angular precision of the FP32 operator: 0 = FP32 arithmetic throughout,
// 1 = BF16 matrix-core projection with FP32 accumulation, 3 = the same with
// the hi/lo split (three BF16 products, ~16 significant bits), 4 = FP32
// products on the FP32 matrix cores where the static class layout applies
// (33..80 harmonics), else as 3.
extern "C" int tr_ctx_set_angular_precision(tr_ctx *c, int prec) {
    if (!c || (prec != 0 && prec != 1 && prec != 3 && prec != 4))
        return fail(TR_ERR_ARG, "tr_ctx_set_angular_precision: 0, 1, 3 or 4");
    c->ang_prec = prec;
    if (c->have32) {
        const HostOperator &op = c->op;
        // the K-blocked schedule's tables hold FP32 or BF16 blocks
        if (v5_ok(op) && c->b5.stat && (c->b5.f32 != 0) != (prec == 4)) {
            const int rc = build_tables5(c);
            if (rc) return rc;
        }
        if (prec != 0 && v4_ok(op))
            c->map32 = mfma_map_ctx(c, 4, 2);
        else if (prec != 0 && v5_ok(op))
            c->map32 = mfma5_map_ctx(c, 8);
        else {
            c->map32 = choose_map(op.slm, 8);
            c->map32.variant = valu_variant(op, 4, c->map32.variant);
            fit_lds(c->map32, op, 4, c->device);
        }
    }
    return TR_OK;
}
